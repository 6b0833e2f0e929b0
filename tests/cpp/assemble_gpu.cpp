// Runs the C++ face of the library (include/anifem_fem3dtet.hpp) with data on
// the GPU, as a reference-side caller would: the element evaluator
// fem3Dtet<GRAD P2, GRAD P2, DfuncTraits<TENSOR_SYMMETRIC, false>> on a batch
// of tets, and Assembler<GRAD P2, GRAD P2>::Assemble of the whole mesh
// (per-tet symmetric tensor, unit load, boundary Dirichlet; homogeneous, then
// with Dirichlet values), the same calls with operators chosen at run time
// (ApplyOp, FemSpace, RuntimeAssembler), and the several-variable context of
// a product space.  Inputs and outputs are raw binary files in the
// directory given on the command line; the test that drives it compares the
// outputs with the oracle.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <string>
#include <vector>

#include "anifem_fem3dtet.hpp"

using namespace anifem_gpu;

namespace {

std::string dir;

template <typename T>
std::vector<T> load(const char* name, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen((dir + "/" + name).c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", name);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename T>
void save(const char* name, const T* data, size_t count) {
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    if (!f || std::fwrite(data, sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot write %s\n", name);
        std::exit(2);
    }
    std::fclose(f);
}

template <typename T>
T* to_device(const std::vector<T>& h) {
    void* p = nullptr;
    if (hipMalloc(&p, h.size() * sizeof(T)) != hipSuccess ||
        hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "device copy failed\n");
        std::exit(3);
    }
    return static_cast<T*>(p);
}

template <typename T>
std::vector<T> to_host(const T* d, size_t count) {
    std::vector<T> h(count);
    if (hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "host copy failed\n");
        std::exit(3);
    }
    return h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) return 1;
    dir = argv[1];
    const auto meta = load<int64_t>("meta.bin", 2);
    const int64_t n_nodes = meta[0], n_tets = meta[1];
    const auto coords = load<double>("coords.bin", 3 * n_nodes);
    const auto tets = load<int32_t>("tets.bin", 4 * n_tets);
    const auto dpack = load<double>("dpacked.bin", 6 * n_tets);  // SoA packed upper triangle
    const auto xyz = load<double>("xyz.bin", 12 * n_tets);       // SoA vertices of the element batch

    using GP2 = Operator<GRAD, FemFix<FEM_P2>>;
    using Traits = DfuncTraits<TENSOR_SYMMETRIC, false>;
    double* d_D = to_device(dpack);
    const TensorView D{TENSOR_SYMMETRIC, ANIFEM_COEF_PER_TET, d_D};

    // element level: the reference's fem3Dtet call shape (int_tet.h:17-20)
    double* d_xyz = to_device(xyz);
    double* d_A = nullptr;
    if (hipMalloc((void**)&d_A, sizeof(double) * 100 * n_tets) != hipSuccess) return 3;
    int rc = fem3Dtet<GP2, GP2, Traits>(d_xyz, n_tets, D, d_A, 2);
    if (rc) {
        std::fprintf(stderr, "fem3Dtet rc=%d\n", rc);
        return 4;
    }
    save("elem.bin", to_host(d_A, 100 * (size_t)n_tets).data(), 100 * (size_t)n_tets);

    // global level: the reference's Assembler call shape (assembler.h:383-386)
    const std::vector<double> one{1.0};
    double* d_f = to_device(one);
    const TensorView F{TENSOR_SCALAR, ANIFEM_COEF_CONST, d_f};
    Assembler<GP2, GP2> asmb(n_nodes, coords.data(), n_tets, tets.data(), 2, D, 2, F, ANIFEM_DIRICHLET_BOUNDARY);
    if (asmb.status()) {
        std::fprintf(stderr, "Assembler status %d\n", asmb.status());
        return 5;
    }
    const int64_t ndof = asmb.ndof(), nnz = asmb.nnz();
    double *d_val = nullptr, *d_rhs = nullptr;
    if (hipMalloc((void**)&d_val, sizeof(double) * nnz) != hipSuccess ||
        hipMalloc((void**)&d_rhs, sizeof(double) * ndof) != hipSuccess)
        return 3;
    anifem_status st{};
    rc = asmb.Assemble(D, &F, d_val, d_rhs, &st);
    if (rc || st.code) {
        std::fprintf(stderr, "Assemble rc=%d status=%d\n", rc, st.code);
        return 6;
    }
    const int64_t* d_rp = nullptr;
    const int32_t* d_ci = nullptr;
    if (asmb.csr(&d_rp, &d_ci)) return 7;
    std::vector<int32_t> rtets(4 * n_tets);
    if (asmb.Tets(rtets.data())) return 7;
    const int64_t sizes[2] = {ndof, nnz};
    save("sizes.bin", sizes, 2);
    save("val.bin", to_host(d_val, nnz).data(), nnz);
    save("rhs.bin", to_host(d_rhs, ndof).data(), ndof);
    save("row_ptr.bin", to_host(d_rp, ndof + 1).data(), ndof + 1);
    save("col_idx.bin", to_host(d_ci, nnz).data(), nnz);
    save("tets_reordered.bin", rtets.data(), rtets.size());

    // Dirichlet values: bc(dof) = 1 + dof / ndof on every DOF (only the
    // Dirichlet ones are read)
    std::vector<double> bc(ndof);
    for (int64_t d = 0; d < ndof; ++d) bc[d] = 1.0 + (double)d / (double)ndof;
    double* d_bc = to_device(bc);
    rc = asmb.Assemble(D, &F, d_val, d_rhs, &st, FP64, nullptr, d_bc);
    if (rc || st.code) return 8;
    save("rhs_bc.bin", to_host(d_rhs, ndof).data(), ndof);

    // operators chosen at run time (int_tet.h:145-147; operators.h:418-533,
    // fem_space.h:404-596): the same calls through ApplyOp / FemSpace give the
    // template calls' results bit for bit
    const FemSpace P2{P2Space{}};
    const ApplyOp g2 = P2.getOP(GRAD);
    double* d_A2 = nullptr;
    if (hipMalloc((void**)&d_A2, sizeof(double) * 100 * n_tets) != hipSuccess) return 3;
    rc = fem3Dtet(d_xyz, n_tets, D, g2, g2, d_A2, 2);
    if (rc) return 9;
    if (to_host(d_A2, 100 * (size_t)n_tets) != to_host(d_A, 100 * (size_t)n_tets)) {
        std::fprintf(stderr, "runtime fem3Dtet differs from the template call\n");
        return 10;
    }
    RuntimeAssembler rasm(g2, P2.getOP(GRAD), n_nodes, coords.data(), n_tets, tets.data(), 2, D, 2, F,
                          ANIFEM_DIRICHLET_BOUNDARY);
    if (rasm.status() || rasm.ndof() != ndof || rasm.nnz() != nnz) return 11;
    double *d_val2 = nullptr, *d_rhs2 = nullptr;
    if (hipMalloc((void**)&d_val2, sizeof(double) * nnz) != hipSuccess ||
        hipMalloc((void**)&d_rhs2, sizeof(double) * ndof) != hipSuccess)
        return 3;
    rc = rasm.Assemble(D, &F, d_val2, d_rhs2, &st);
    if (rc || st.code) return 12;
    rc = asmb.Assemble(D, &F, d_val, d_rhs, &st);
    if (rc || st.code) return 12;
    if (to_host(d_val2, nnz) != to_host(d_val, nnz) || to_host(d_rhs2, ndof) != to_host(d_rhs, ndof)) {
        std::fprintf(stderr, "runtime assembler differs from the template one\n");
        return 13;
    }
    // a product space (mixed Poisson: RT0 x P0) builds the several-variable
    // context: one DOF per face, then one per cell
    const FemSpace mixed = FemSpace(RT0Space{}) * FemSpace(P0Space{});
    RuntimeAssembler masm(mixed, n_nodes, coords.data(), n_tets, tets.data(), ANIFEM_DIRICHLET_NONE);
    if (masm.status()) return 14;
    const int64_t mixed_sizes[2] = {masm.ndof(), masm.nnz()};
    save("mixed_sizes.bin", mixed_sizes, 2);
    std::printf("cpp face ok: %lld dofs, %lld nnz\n", (long long)ndof, (long long)nnz);
    return 0;
}
