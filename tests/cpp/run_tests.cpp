// Tests of the C++ SMAX::Interface with the CPU and HIP platforms.
//
//   run_tests host   API behaviour that needs no GPU (registration, argument
//                    checks, refused platforms and types), and the
//                    reference's known answers through PlatformType::CPU
//   run_tests gpu    the reference's SpMV / SELL / BCRS / SpMM known answers
//                    (examples/tests/*.hpp) run through PlatformType::HIP,
//                    the Jacobi application loop against a host Jacobi, and
//                    the CPU and HIP platforms against each other bit for bit
//
// Exit status 0 when every selected test passes.
#include <cmath>
#include <cstring>
#include <limits>
#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "interface.hpp"

using SMAX::FloatType;
using SMAX::IntType;
using SMAX::KernelType;
using SMAX::PlatformType;

namespace {

std::map<std::string, std::function<void()>> &registry(const std::string &group) {
    static std::map<std::string, std::map<std::string, std::function<void()>>> r;
    return r[group];
}

struct Reg {
    Reg(const char *group, const char *name, std::function<void()> f) {
        registry(group)[name] = std::move(f);
    }
};

#define TEST(group, name)                                                      \
    static void name();                                                        \
    static Reg reg_##name(#group, #name, name);                                \
    static void name()

void check(bool ok, const std::string &what) {
    if (!ok) throw std::runtime_error(what);
}

// |a-b| <= 1e-14, else relative 4 eps (examples/tests/tests_common.hpp:9-35)
template <typename T> void close(const T *a, const T *b, size_t n, const char *what) {
    const T abs_eps = sizeof(T) == 8 ? T(1e-14) : T(1e-7);
    const T rel = std::numeric_limits<T>::epsilon() * 4;
    for (size_t i = 0; i < n; ++i) {
        const T d = std::fabs(a[i] - b[i]);
        if (d <= abs_eps || d <= rel * std::max(std::fabs(a[i]), std::fabs(b[i]))) continue;
        throw std::runtime_error(std::string(what) + " mismatch at " + std::to_string(i) +
                                 ": " + std::to_string(a[i]) + " vs " + std::to_string(b[i]));
    }
}

template <typename F> void expect_throw(F f, const char *what) {
    try {
        f();
    } catch (const std::exception &) {
        return;
    }
    throw std::runtime_error(std::string("expected an exception: ") + what);
}

// The 3x3 matrix of examples/tests/spmv_test.hpp:14-17.
std::vector<int> k_col = {0, 1, 1, 0, 2};
std::vector<int> k_rp = {0, 2, 3, 5};
std::vector<double> k_val = {1.1, 1.2, 2.2, 3.1, 3.3};

} // namespace

// ------------------------------------------------------------------- host
TEST(host, refuses_platforms_not_built) {
    SMAX::Interface smax;
    expect_throw([&] { smax.register_kernel("b", KernelType::SPMV, PlatformType::CUDA); },
                 "CUDA platform");
    // the reference's default platform (interface.hpp:33-36) and HIP
    check(smax.register_kernel("a", KernelType::SPMV) == 0, "CPU (default)");
    check(smax.register_kernel("c", KernelType::SPMV, PlatformType::HIP) == 0, "HIP");
}

TEST(host, short_operands_are_refused) {
    // x or y smaller than the product reads / writes (n_cols / n_rows entries
    // past the element offset): run() and apply() report failure on either
    // platform before touching memory or the GPU
    for (auto p : {PlatformType::CPU, PlatformType::HIP}) {
        SMAX::Interface smax;
        std::vector<double> x(3, 1.0), y(3, 0.0);
        smax.register_kernel("k", KernelType::SPMV, p);
        auto *k = smax.kernel("k");
        k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
        k->register_B(2, x.data());
        k->register_C(3, y.data());
        check(k->run() == 1, "x of 2 entries for 3 columns");
        k->register_B(3, x.data());
        k->register_C(2, y.data());
        check(k->run() == 1, "y of 2 entries for 3 rows");
        k->register_C(3, y.data());
        check(k->apply(0, 1, 0) == 1 && k->apply(0, 0, 1) == 1, "offsets past the end");
        check(y[0] == 0.0 && y[1] == 0.0 && y[2] == 0.0, "y untouched");
    }
}

// ------------------------------------------------------- host: CPU platform
TEST(host, cpu_spmv_crs_kat) {  // examples/tests/spmv_test.hpp, default platform
    SMAX::Interface smax;
    std::vector<double> x(3, 1.0), y(3), want = {2.3, 2.2, 6.4};
    smax.register_kernel("spmv", KernelType::SPMV);
    auto *k = smax.kernel("spmv");
    k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
    k->register_B(3, x.data());
    k->register_C(3, y.data());
    check(k->run() == 0, "run");
    close(y.data(), want.data(), 3, "spmv_y");
}

template <typename IT> void cpu_crs_index_type(IntType it) {
    SMAX::Interface smax;
    std::vector<IT> col(k_col.begin(), k_col.end()), rp(k_rp.begin(), k_rp.end());
    std::vector<double> x(3, 1.0), y(3), want = {2.3, 2.2, 6.4};
    smax.register_kernel("k", KernelType::SPMV, PlatformType::CPU, it);
    auto *k = smax.kernel("k");
    k->register_A(3, 3, 5, col.data(), rp.data(), k_val.data());
    k->register_B(3, x.data());
    k->register_C(3, y.data());
    check(k->run() == 0, "run");
    close(y.data(), want.data(), 3, "index type");
}

TEST(host, cpu_every_index_type) {  // the CPU dispatch covers all six (spmv_cpu.hpp:38-96)
    cpu_crs_index_type<int16_t>(IntType::INT16);
    cpu_crs_index_type<int32_t>(IntType::INT32);
    cpu_crs_index_type<int64_t>(IntType::INT64);
    cpu_crs_index_type<uint16_t>(IntType::UINT16);
    cpu_crs_index_type<uint32_t>(IntType::UINT32);
    cpu_crs_index_type<uint64_t>(IntType::UINT64);
}

TEST(host, cpu_spmv_scs_kat) {
    // spmv_scs_2_4_test_1.hpp: C = 2, sigma = 4: rows {0, 2} form chunk 0
    // (2 entries each), row 1 and a padding slot chunk 1; y in SELL order
    SMAX::Interface smax;
    std::vector<int> chunk_ptr = {0, 4, 6}, chunk_len = {2, 1}, col = {0, 0, 1, 2, 1, 0},
                     perm = {0, 2, 1};
    std::vector<double> val = {1.1, 3.1, 1.2, 3.3, 2.2, 0.0};
    std::vector<double> x(3, 1.0), y(4, -1.0), want = {2.3, 6.4, 2.2, 0.0};
    smax.register_kernel("scs", KernelType::SPMV, PlatformType::CPU);
    auto *k = smax.kernel("scs");
    k->set_mat_scs(true);
    k->register_A(2, 4, 3, 4, 3, 2, 6, 5, chunk_ptr.data(), chunk_len.data(), col.data(),
                  val.data(), perm.data());
    k->register_B(3, x.data());
    k->register_C(4, y.data());
    check(k->run() == 0, "run");
    close(y.data(), want.data(), 4, "scs_y");
}

TEST(host, cpu_spmv_scs_original_order) {
    // the SELL KAT above with y in original row order: y[r] = y_scs[perm[r]]
    SMAX::Interface smax;
    std::vector<int> chunk_ptr = {0, 4, 6}, chunk_len = {2, 1}, col = {0, 0, 1, 2, 1, 0},
                     perm = {0, 2, 1};
    std::vector<double> val = {1.1, 3.1, 1.2, 3.3, 2.2, 0.0};
    std::vector<double> x(3, 1.0), y(3, -1.0), want = {2.3, 2.2, 6.4};
    smax.register_kernel("scs", KernelType::SPMV, PlatformType::CPU);
    auto *k = smax.kernel("scs");
    k->set_mat_scs(true);
    check(k->set_scs_original_order(true) == 0, "flag");
    k->register_A(2, 4, 3, 4, 3, 2, 6, 5, chunk_ptr.data(), chunk_len.data(), col.data(),
                  val.data(), perm.data());
    k->register_B(3, x.data());
    k->register_C(3, y.data());  // n_rows entries suffice in original order
    check(k->run() == 0, "run");
    close(y.data(), want.data(), 3, "scs_y original order");
}

TEST(host, cpu_bcrs_kats) {  // spmv_bcrs_1_1_test_3.hpp, both block layouts
    std::vector<int> col = {0, 1, 1}, rp = {0, 2, 3};
    // block (ix, jx): column-major at jx*2 + ix, row-major at ix*2 + jx
    std::vector<double> cm = {1.1, 0.0, 0.0, 2.2, 1.3, 0.0, 1.4, 0.0, 0.0, 0.0, 0.0, 4.4};
    std::vector<double> rm = {1.1, 0.0, 0.0, 2.2, 1.3, 1.4, 0.0, 0.0, 0.0, 0.0, 0.0, 4.4};
    for (bool col_major : {true, false}) {
        SMAX::Interface smax;
        std::vector<double> x(4, 1.0), y(4), want = {3.8, 2.2, 0.0, 4.4};
        smax.register_kernel("b", KernelType::SPMV, PlatformType::CPU);
        auto *k = smax.kernel("b");
        k->set_mat_bcrs(true);
        k->register_A(2, 2, 3, 2, 2, 2, 2, col.data(), rp.data(),
                      col_major ? cm.data() : rm.data());
        k->register_B(4, x.data());
        k->register_C(4, y.data());
        k->set_block_column_major(col_major);
        check(k->run() == 0, "run");
        close(y.data(), want.data(), 4, col_major ? "bcrs cm" : "bcrs rm");
    }
}

TEST(host, cpu_spmm_kats) {  // spmm_test.hpp (column-major), spmm_rowmaj_test.hpp
    for (bool row_major : {false, true}) {
        SMAX::Interface smax;
        const int k = 4;
        std::vector<double> X(3 * k, 1.0), Y(3 * k);
        std::vector<double> want = {2.3, 2.2, 6.4, 2.3, 2.2, 6.4, 2.3, 2.2, 6.4, 2.3, 2.2, 6.4};
        if (row_major) want = {2.3, 2.3, 2.3, 2.3, 2.2, 2.2, 2.2, 2.2, 6.4, 6.4, 6.4, 6.4};
        smax.register_kernel("spmm", KernelType::SPMM);
        auto *kk = smax.kernel("spmm");
        kk->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
        kk->register_B(3, k, X.data());
        kk->register_C(3, k, Y.data());
        if (row_major) kk->set_vec_row_major(true);
        check(kk->run() == 0, "run");
        close(Y.data(), want.data(), 3 * k, "spmm_Y");
    }
}

TEST(host, cpu_swap_operands_and_offsets) {
    // x_{k+1} = A x_k three times through swap_operands, then one product at
    // element offsets 1 (spmv_cpu_core.hpp:62-63): the same fma chains as a
    // host loop, bit for bit
    SMAX::Interface smax;
    std::vector<double> a(3, 1.0), b(3, 0.0), host(3, 1.0), tmp(3);
    smax.register_kernel("it", KernelType::SPMV, PlatformType::CPU);
    auto *k = smax.kernel("it");
    k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
    k->register_B(3, a.data());
    k->register_C(3, b.data());
    for (int it = 0; it < 3; ++it) {
        check(k->run() == 0, "run");
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int j = k_rp[r]; j < k_rp[r + 1]; ++j) s = std::fma(k_val[j], host[k_col[j]], s);
            tmp[r] = s;
        }
        host = tmp;
        k->swap_operands();
    }
    check(std::memcmp(b.data(), host.data(), sizeof(double) * 3) == 0, "swapped iterate");
    std::vector<double> xl = {9.0, 1.0, 2.0, 3.0}, yl(4, 0.0);
    smax.register_kernel("off", KernelType::SPMV);
    auto *o = smax.kernel("off");
    o->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
    o->register_B(4, xl.data());
    o->register_C(4, yl.data());
    check(o->initialize() == 0 && o->apply(0, 1, 1) == 0 && o->finalize() == 0, "offsets");
    const double w1 = std::fma(1.2, 2.0, 1.1 * 1.0);
    check(yl[0] == 0.0 && yl[1] == w1, "offset y");
}

TEST(host, missing_kernel_throws) {
    SMAX::Interface smax;
    expect_throw([&] { smax.kernel("nope"); }, "kernel_dne");
}

TEST(host, argument_counts_and_metadata_are_checked) {
    SMAX::Interface smax;
    smax.register_kernel("k", KernelType::SPMV, PlatformType::HIP);
    auto *k = smax.kernel("k");
    expect_throw([&] { k->register_A(3, 3, k_col.data(), k_rp.data(), k_val.data()); },
                 "CRS with 5 args");
    expect_throw([&] { k->register_A(-3, 3, 5, k_col.data(), k_rp.data(), k_val.data()); },
                 "negative metadata");
    expect_throw([&] { k->register_B(3); }, "register_B with 1 arg");
    k->set_mat_scs(true);
    expect_throw([&] { k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data()); },
                 "SCS with 6 args");
    smax.register_kernel("m", KernelType::SPMM, PlatformType::HIP);
    check(smax.kernel("m")->set_mat_scs(true) == 1, "SpMM has no SCS flag");
    check(smax.kernel("m")->set_vec_row_major(true) == 0, "SpMM row-major flag");
}

TEST(host, unsupported_index_type_fails_run) {
    SMAX::Interface smax;
    smax.register_kernel("k", KernelType::SPMV, PlatformType::HIP, IntType::INT64);
    std::vector<long long> col = {0, 1, 1, 0, 2}, rp = {0, 2, 3, 5};
    std::vector<double> x(3, 1.0), y(3);
    auto *k = smax.kernel("k");
    k->register_A(3, 3, 5, col.data(), rp.data(), k_val.data());
    k->register_B(3, x.data());
    k->register_C(3, y.data());
    check(k->run() == 1, "INT64 indices are refused on HIP");
}

// -------------------------------------------------------------------- gpu
TEST(gpu, spmv_crs_kat) {  // examples/tests/spmv_test.hpp
    SMAX::Interface smax;
    std::vector<double> x(3, 1.0), y(3), want = {2.3, 2.2, 6.4};
    for (auto impl : {SMAX::SpMVType::naive_thread_per_row, SMAX::SpMVType::naive_warp_group}) {
        smax.register_kernel("spmv", KernelType::SPMV, PlatformType::HIP);
        auto *k = smax.kernel("spmv");
        k->set_kernel_implementation(impl);
        k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
        k->register_B(3, x.data());
        k->register_C(3, y.data());
        check(k->run() == 0, "run");
        close(y.data(), want.data(), 3, "spmv_y");
    }
}

TEST(gpu, spmv_scs_via_gpu_conversion) {
    // spmv_scs_2_4_test_1.hpp: C=2, sigma=4 -> perm {0,2,1}; y_scs[perm] == y_crs
    SMAX::Interface smax;
    unsigned long long C = 2, sigma = 4, n_rows = 0, n_rows_padded = 0, n_cols = 0,
                       n_chunks = 0, n_elements = 0, nnz = 0;
    int *chunk_ptr = nullptr, *chunk_lengths = nullptr, *col = nullptr, *perm = nullptr;
    double *val = nullptr;
    smax.utils->convert_crs_to_scs<int, double, unsigned long long>(
        3, 3, 5, k_col.data(), k_rp.data(), k_val.data(), C, sigma, n_rows,
        n_rows_padded, n_cols, n_chunks, n_elements, nnz, chunk_ptr, chunk_lengths,
        col, val, perm);
    check(n_elements == 6 && n_chunks == 2 && n_rows_padded == 4, "metadata");
    check(perm[0] == 0 && perm[1] == 2 && perm[2] == 1, "perm");
    std::vector<double> x(3, 1.0), y(n_rows_padded), want = {2.3, 2.2, 6.4};
    smax.register_kernel("scs", KernelType::SPMV, PlatformType::HIP);
    auto *k = smax.kernel("scs");
    k->set_mat_scs(true);
    k->register_A(C, sigma, n_rows, n_rows_padded, n_cols, n_chunks, n_elements, nnz,
                  chunk_ptr, chunk_lengths, col, val, perm);
    k->register_B(n_cols, x.data());
    k->register_C(n_rows_padded, y.data());
    check(k->run() == 0, "run");
    for (int i = 0; i < 3; ++i) close(&y[perm[i]], &want[i], 1, "y_scs[perm]");
    delete[] chunk_ptr;
    delete[] chunk_lengths;
    delete[] col;
    delete[] val;
    delete[] perm;
}

TEST(gpu, bcrs_kat) {  // spmv_bcrs_1_1_test_3.hpp
    SMAX::Interface smax;
    std::vector<int> col = {0, 1, 1}, rp = {0, 2, 3};
    std::vector<double> val = {1.1, 0.0, 0.0, 2.2, 1.3, 0.0, 1.4, 0.0, 0.0, 0.0, 0.0, 4.4};
    std::vector<double> x(4, 1.0), y(4), want = {3.8, 2.2, 0.0, 4.4};
    smax.register_kernel("b", KernelType::SPMV, PlatformType::HIP);
    auto *k = smax.kernel("b");
    k->set_mat_bcrs(true);
    k->register_A(2, 2, 3, 2, 2, 2, 2, col.data(), rp.data(), val.data());
    k->register_B(4, x.data());
    k->register_C(4, y.data());
    k->set_block_column_major(true);
    check(k->run() == 0, "run");
    close(y.data(), want.data(), 4, "bcrs_y");
}

TEST(gpu, bcrs_laplace_equals_crs) {  // spmv_bcrs_laplace_test.hpp, 4x8 blocks
    const int n = 16;
    std::vector<int> rp(1, 0), col;
    std::vector<double> val;
    for (int i = 0; i < n; ++i) {
        for (int j = i - 1; j <= i + 1; ++j)
            if (j >= 0 && j < n) {
                col.push_back(j);
                val.push_back(i == j ? 2.0 : -1.0);
            }
        rp.push_back(static_cast<int>(col.size()));
    }
    std::vector<double> x(n), y_crs(n), y_b(n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 / double((i + 3) % 23 + 1);
    for (bool cm : {true, false}) {
        SMAX::Interface smax;
        unsigned long long nbr = 0, nbc = 0, nb = 0, bh = 0, bw = 0, hp = 0, wp = 0;
        int *bcol = nullptr, *brp = nullptr;
        double *bval = nullptr;
        smax.utils->convert_crs_to_bcrs<int, double, unsigned long long>(
            n, n, col.size(), col.data(), rp.data(), val.data(), nbr, nbc, nb, bh, bw,
            hp, wp, bcol, brp, bval, 4, 8, 4, 8, cm);
        smax.register_kernel("crs", KernelType::SPMV, PlatformType::HIP);
        smax.kernel("crs")->register_A(n, n, col.size(), col.data(), rp.data(), val.data());
        smax.kernel("crs")->register_B(n, x.data());
        smax.kernel("crs")->register_C(n, y_crs.data());
        check(smax.kernel("crs")->run() == 0, "crs run");
        smax.register_kernel("b", KernelType::SPMV, PlatformType::HIP);
        auto *k = smax.kernel("b");
        k->set_mat_bcrs(true);
        k->register_A(nbr, nbc, nb, bh, bw, hp, wp, bcol, brp, bval);
        k->register_B(nbc * wp, x.data());
        k->register_C(nbr * hp, y_b.data());
        k->set_block_column_major(cm);
        check(k->run() == 0, "bcrs run");
        close(y_b.data(), y_crs.data(), n, "laplace");
        delete[] bcol;
        delete[] brp;
        delete[] bval;
    }
}

TEST(gpu, bcrs_conversion_rejects_bad_blocks) {
    SMAX::Interface smax;
    unsigned long long a = 0, b = 0, c = 0, d = 0, e = 0, f = 0, g = 0;
    int *bc = nullptr, *br = nullptr;
    double *bv = nullptr;
    expect_throw([&] {
        smax.utils->convert_crs_to_bcrs<int, double, unsigned long long>(
            3, 3, 5, k_col.data(), k_rp.data(), k_val.data(), a, b, c, d, e, f, g, bc,
            br, bv, 2, 2, 2, 2, true);
    }, "block height must divide n_rows");
}

TEST(gpu, spmm_kats) {  // spmm_test.hpp (column-major), spmm_rowmaj_test.hpp
    for (bool row_major : {false, true}) {
        SMAX::Interface smax;
        const int k = 4;
        std::vector<double> X(3 * k, 1.0), Y(3 * k);
        std::vector<double> want = {2.3, 2.2, 6.4, 2.3, 2.2, 6.4, 2.3, 2.2, 6.4, 2.3, 2.2, 6.4};
        if (row_major) want = {2.3, 2.3, 2.3, 2.3, 2.2, 2.2, 2.2, 2.2, 6.4, 6.4, 6.4, 6.4};
        smax.register_kernel("spmm", KernelType::SPMM, PlatformType::HIP);
        auto *kk = smax.kernel("spmm");
        kk->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
        kk->register_B(3, k, X.data());
        kk->register_C(3, k, Y.data());
        if (row_major) kk->set_vec_row_major(true);
        check(kk->run() == 0, "run");
        close(Y.data(), want.data(), 3 * k, "spmm_Y");
    }
}

TEST(gpu, spmm_column_major_refuses_overrun) {
    // 5 x 3 matrix, column-major with the reference's leading dimension
    // n_rows = 5: X (3 x k) would be read up to (k-1)*5 + 3 > 3k entries, so
    // run() reports failure instead of reading past the registered buffer
    SMAX::Interface smax;
    const int k = 2;
    std::vector<int> rp = {0, 1, 2, 3, 3, 3}, col = {0, 1, 2};
    std::vector<double> val = {1.0, 2.0, 3.0}, X(3 * k, 1.0), Y(5 * k, 0.0);
    smax.register_kernel("spmm_overrun", KernelType::SPMM, PlatformType::HIP);
    auto *kk = smax.kernel("spmm_overrun");
    kk->register_A(5, 3, 3, col.data(), rp.data(), val.data());
    kk->register_B(3, k, X.data());
    kk->register_C(5, k, Y.data());
    check(kk->run() != 0, "overrun refused");
}

TEST(gpu, float32_and_bfloat16_values) {
    // 2-D 5-point Laplacian on 32^2 with small-integer x: exact in every type
    const int n = 32, N = n * n;
    std::vector<int> rp(1, 0), col;
    std::vector<double> v64;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int r = i * n + j;
            if (j > 0) { col.push_back(r - 1); v64.push_back(-1); }
            if (i > 0) { col.push_back(r - n); v64.push_back(-1); }
            col.push_back(r); v64.push_back(4);
            if (i < n - 1) { col.push_back(r + n); v64.push_back(-1); }
            if (j < n - 1) { col.push_back(r + 1); v64.push_back(-1); }
            rp.push_back(static_cast<int>(col.size()));
        }
    std::vector<float> x(N), y(N), want(N), v32(v64.begin(), v64.end());
    for (int i = 0; i < N; ++i) x[i] = float(i % 7) - 3.0f;
    for (int r = 0; r < N; ++r) {
        float s = 0;
        for (int j = rp[r]; j < rp[r + 1]; ++j) s += v32[j] * x[col[j]];
        want[r] = s;
    }
    std::vector<uint16_t> vbf(v64.size());
    for (size_t i = 0; i < v64.size(); ++i) {  // -1 and 4 are exact in bfloat16
        float f = static_cast<float>(v64[i]);
        uint32_t u;
        std::memcpy(&u, &f, 4);
        vbf[i] = static_cast<uint16_t>(u >> 16);
    }
    SMAX::Interface smax;
    smax.register_kernel("f32", KernelType::SPMV, PlatformType::HIP, IntType::INT32,
                         FloatType::FLOAT32);
    auto *k = smax.kernel("f32");
    k->register_A(N, N, col.size(), col.data(), rp.data(), v32.data());
    k->register_B(N, x.data());
    k->register_C(N, y.data());
    check(k->run() == 0, "f32 run");
    close(y.data(), want.data(), N, "f32");
    std::fill(y.begin(), y.end(), 0.0f);
    smax.register_kernel("bf16", KernelType::SPMV, PlatformType::HIP, IntType::INT32,
                         FloatType::BFLOAT16);
    k = smax.kernel("bf16");
    k->register_A(N, N, col.size(), col.data(), rp.data(), vbf.data());
    k->register_B(N, x.data());
    k->register_C(N, y.data());
    check(k->run() == 0, "bf16 run");
    close(y.data(), want.data(), N, "bf16");
}

TEST(gpu, jacobi_style_swap_and_offsets) {
    // x_{k+1} = A x_k three times through swap_operands (jacobi.cpp:26-58
    // pattern), then one product into y at an offset.
    SMAX::Interface smax;
    std::vector<double> a(3, 1.0), b(3, 0.0), host(3, 1.0), tmp(3);
    smax.register_kernel("it", KernelType::SPMV, PlatformType::HIP);
    auto *k = smax.kernel("it");
    k->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
    k->register_B(3, a.data());
    k->register_C(3, b.data());
    for (int it = 0; it < 3; ++it) {
        check(k->run() == 0, "run");
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int j = k_rp[r]; j < k_rp[r + 1]; ++j) s = std::fma(k_val[j], host[k_col[j]], s);
            tmp[r] = s;
        }
        host = tmp;
        k->swap_operands();
    }
    // iterations 1 and 3 wrote b, iteration 2 wrote a
    close(b.data(), host.data(), 3, "swapped iterate");
    // one product with x read at offset 1 and y written at offset 1
    std::vector<double> xl = {9.0, 1.0, 2.0, 3.0}, yl(4, 0.0);
    smax.register_kernel("off", KernelType::SPMV, PlatformType::HIP);
    auto *o = smax.kernel("off");
    o->register_A(3, 3, 5, k_col.data(), k_rp.data(), k_val.data());
    o->register_B(4, xl.data());
    o->register_C(4, yl.data());
    check(o->initialize() == 0 && o->apply(0, 1, 1) == 0, "offset apply");
    check(o->finalize() == 0, "offset finalize");
    // finalize copies n_rows entries from the start of device y
    // (spmv_cuda_core.cuh:190-200): y[1..3] = A * (1,2,3), y[0] untouched
    const double w1 = std::fma(1.2, 2.0, 1.1 * 1.0);
    close(&yl[1], &w1, 1, "offset y[1]");
}

TEST(gpu, jacobi_poisson_matches_host_iterates) {
    // The reference's Jacobi application loop (examples/applications/jacobi/
    // jacobi.cpp:26-58): x_new <- A x_old through run(), the host normalises
    // x_new = (b - (A x_old - D x_old)) / D, then swap_operands and the
    // application's pointer swap.  Every run re-uploads the swapped
    // operands.  The GPU SpMV is the reference fma chain, so the iterates
    // equal a host Jacobi with the same chain bit for bit, iteration by
    // iteration, until the residual test stops both.
    const int n = 24, N = n * n;
    std::vector<int> rp(1, 0), cl;
    std::vector<double> va;
    for (int r = 0; r < N; ++r) {
        const int i = r / n, j = r % n;
        if (i > 0) { cl.push_back(r - n); va.push_back(-1.0); }
        if (j > 0) { cl.push_back(r - 1); va.push_back(-1.0); }
        cl.push_back(r); va.push_back(4.0);
        if (j < n - 1) { cl.push_back(r + 1); va.push_back(-1.0); }
        if (i < n - 1) { cl.push_back(r + n); va.push_back(-1.0); }
        rp.push_back(static_cast<int>(cl.size()));
    }
    const std::vector<double> b(N, 1.0), D(N, 4.0);
    auto spmv_host = [&](const std::vector<double> &x, std::vector<double> &y) {
        for (int r = 0; r < N; ++r) {
            double s = 0;
            for (int e = rp[r]; e < rp[r + 1]; ++e) s = std::fma(va[e], x[cl[e]], s);
            y[r] = s;
        }
    };
    auto normalize = [&](std::vector<double> &xn, const std::vector<double> &xo) {
        for (int r = 0; r < N; ++r) xn[r] = (b[r] - (xn[r] - D[r] * xo[r])) / D[r];
    };
    auto residual = [&](const std::vector<double> &x) {
        std::vector<double> t(N);
        spmv_host(x, t);
        double m = 0;
        for (int r = 0; r < N; ++r) m = std::max(m, std::fabs(b[r] - t[r]));
        return m;
    };
    const double tol = 1e-6;
    const int max_iters = 5000, check_every = 10;
    // host Jacobi
    std::vector<double> ho(N, 0.0), hn(N, 0.0);
    int host_iters = 0;
    for (double res = residual(ho); res > tol && host_iters < max_iters;) {
        spmv_host(ho, hn);
        normalize(hn, ho);
        std::swap(ho, hn);
        if (++host_iters % check_every == 0) res = residual(ho);
    }
    // the same loop with the product on the GPU through SMAX::Interface
    std::vector<double> a(N, 0.0), c(N, 0.0);
    std::vector<double> *x_old = &a, *x_new = &c;
    SMAX::Interface smax;
    smax.register_kernel("x_new <- Ax_old", KernelType::SPMV, PlatformType::HIP);
    auto *k = smax.kernel("x_new <- Ax_old");
    k->register_A(N, N, static_cast<int>(cl.size()), cl.data(), rp.data(), va.data());
    k->register_B(N, x_old->data());
    k->register_C(N, x_new->data());
    int iters = 0;
    for (double res = residual(*x_old); res > tol && iters < max_iters;) {
        check(k->run() == 0, "jacobi run");
        normalize(*x_new, *x_old);
        k->swap_operands();
        std::swap(x_old, x_new);
        if (++iters % check_every == 0) res = residual(*x_old);
    }
    check(iters == host_iters && iters < max_iters,
          "iteration count " + std::to_string(iters) + " vs host " + std::to_string(host_iters));
    check(std::memcmp(x_old->data(), ho.data(), sizeof(double) * N) == 0,
          "GPU-driven iterate differs from the host Jacobi");
}

namespace {

// A 3000-row band matrix with 1..40 entries per row (sorted, distinct
// columns), values and x from a fixed LCG: the cross-platform check input.
struct Band {
    int n = 3000;
    std::vector<int> rp{0}, col;
    std::vector<double> val, x;
    Band() {
        uint64_t s = 12345;
        auto next = [&] { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return s >> 11; };
        for (int r = 0; r < n; ++r) {
            const int len = 1 + static_cast<int>(next() % 40);
            std::vector<int> c;
            while (static_cast<int>(c.size()) < len) {
                const int j = std::min(n - 1, std::max(0, r - 200 + static_cast<int>(next() % 401)));
                if (std::find(c.begin(), c.end(), j) == c.end()) c.push_back(j);
            }
            std::sort(c.begin(), c.end());
            for (int j : c) {
                col.push_back(j);
                val.push_back(double(next()) / double(1ULL << 53) * 2 - 1);
            }
            rp.push_back(static_cast<int>(col.size()));
        }
        for (int i = 0; i < n; ++i) x.push_back(double(next()) / double(1ULL << 53) * 2 - 1);
    }
};

template <typename V, typename A>
std::vector<A> run_spmv(PlatformType p, FloatType ft, const Band &m, const std::vector<V> &val,
                        const std::vector<A> &x, SMAX::SpMVType impl) {
    SMAX::Interface smax;
    std::vector<A> xx(x), y(m.n, A(-7));
    smax.register_kernel("k", KernelType::SPMV, p, IntType::INT32, ft);
    auto *k = smax.kernel("k");
    k->set_kernel_implementation(impl);
    k->register_A(m.n, m.n, static_cast<int>(m.col.size()), const_cast<int *>(m.col.data()),
                  const_cast<int *>(m.rp.data()), const_cast<V *>(val.data()));
    k->register_B(m.n, xx.data());
    k->register_C(m.n, y.data());
    check(k->run() == 0, "run");
    return y;
}

} // namespace

TEST(gpu, cpu_and_hip_platforms_agree_bitwise) {
    // The same registered arrays through PlatformType::CPU and ::HIP: CRS
    // SpMV (both HIP schedules) in FP64, FP32 and BF16 values, and CRS SpMM
    // in both layouts, equal bit for bit.
    const Band m;
    std::vector<float> v32(m.val.begin(), m.val.end()), x32(m.x.begin(), m.x.end());
    std::vector<uint16_t> vbf(m.val.size());
    for (size_t i = 0; i < vbf.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &v32[i], 4);
        vbf[i] = static_cast<uint16_t>(u >> 16);  // truncation: any bfloat16 pattern will do
    }
    for (auto impl : {SMAX::SpMVType::naive_thread_per_row, SMAX::SpMVType::naive_warp_group}) {
        auto a = run_spmv(PlatformType::CPU, FloatType::FLOAT64, m, m.val, m.x, impl);
        auto b = run_spmv(PlatformType::HIP, FloatType::FLOAT64, m, m.val, m.x, impl);
        check(std::memcmp(a.data(), b.data(), sizeof(double) * m.n) == 0, "FP64 CRS");
        auto c = run_spmv(PlatformType::CPU, FloatType::FLOAT32, m, v32, x32, impl);
        auto d = run_spmv(PlatformType::HIP, FloatType::FLOAT32, m, v32, x32, impl);
        check(std::memcmp(c.data(), d.data(), sizeof(float) * m.n) == 0, "FP32 CRS");
        auto e = run_spmv(PlatformType::CPU, FloatType::BFLOAT16, m, vbf, x32, impl);
        auto f = run_spmv(PlatformType::HIP, FloatType::BFLOAT16, m, vbf, x32, impl);
        check(std::memcmp(e.data(), f.data(), sizeof(float) * m.n) == 0, "BF16 CRS");
    }
    for (bool row_major : {false, true}) {
        const int k = 8;
        std::vector<double> X(m.n * k), Y[2];
        for (int i = 0; i < m.n * k; ++i) X[i] = m.x[i % m.n] * double(1 + i % 3);
        for (int p = 0; p < 2; ++p) {
            SMAX::Interface smax;
            Y[p].assign(m.n * k, -7.0);
            smax.register_kernel("mm", KernelType::SPMM, p ? PlatformType::HIP : PlatformType::CPU);
            auto *kk = smax.kernel("mm");
            kk->register_A(m.n, m.n, static_cast<int>(m.col.size()),
                           const_cast<int *>(m.col.data()), const_cast<int *>(m.rp.data()),
                           const_cast<double *>(m.val.data()));
            kk->register_B(m.n, k, X.data());
            kk->register_C(m.n, k, Y[p].data());
            if (row_major) kk->set_vec_row_major(true);
            check(kk->run() == 0, "spmm run");
        }
        check(std::memcmp(Y[0].data(), Y[1].data(), sizeof(double) * m.n * k) == 0,
              row_major ? "SpMM row-major" : "SpMM column-major");
    }
}

TEST(gpu, sell_cpu_and_hip_platforms_agree_bitwise) {
    // A registered SELL matrix through PlatformType::CPU (the reference's full
    // padded chunk chains) and ::HIP (a handle over the arrays: slot lengths,
    // schedule and, for the unsorted layouts padded more than 6x by every
    // 61st row's 500 entries, tail waves): the rows of y in SELL order, equal
    // bit for bit (NaN for NaN), for every (C, sigma) and with x[0] = -0.0,
    // +inf and NaN.
    Band m;
    {  // every 61st row gets 500 entries
        std::vector<int> rp{0}, col;
        std::vector<double> val;
        for (int r = 0; r < m.n; ++r) {
            const int len = r % 61 == 0 ? 500 : m.rp[r + 1] - m.rp[r];
            for (int j = 0; j < len; ++j) {
                col.push_back(r % 61 == 0 ? (r + 7 * j) % m.n : m.col[m.rp[r] + j]);
                val.push_back(r % 61 == 0 ? 0.001 * (j + 1) - 0.2 : m.val[m.rp[r] + j]);
            }
            if (r % 61 == 0) std::sort(col.end() - len, col.end());
            rp.push_back(static_cast<int>(col.size()));
        }
        m.rp = rp;
        m.col = col;
        m.val = val;
    }
    for (double x0 : std::initializer_list<double>{0.5, -0.0,
                                                   std::numeric_limits<double>::infinity(),
                                                   std::numeric_limits<double>::quiet_NaN()}) {
        m.x[0] = x0;
        for (auto cs : {std::pair<int, int>{64, 1}, {32, 64}, {64, 4096}, {8, 1}}) {
            SMAX::Interface conv;
            unsigned long long n_rows = 0, n_rows_padded = 0, n_cols = 0, n_chunks = 0,
                               n_elements = 0, nnz = 0;
            int *chunk_ptr = nullptr, *chunk_lengths = nullptr, *col = nullptr, *perm = nullptr;
            double *val = nullptr;
            conv.utils->convert_crs_to_scs<int, double, unsigned long long>(
                m.n, m.n, static_cast<int>(m.col.size()), m.col.data(), m.rp.data(),
                m.val.data(), cs.first, cs.second, n_rows, n_rows_padded, n_cols, n_chunks,
                n_elements, nnz, chunk_ptr, chunk_lengths, col, val, perm);
            std::vector<double> y[2];
            for (int p = 0; p < 2; ++p) {
                SMAX::Interface smax;
                std::vector<double> x(m.x);
                y[p].assign(n_rows_padded, -7.0);
                smax.register_kernel("s", KernelType::SPMV, p ? PlatformType::HIP : PlatformType::CPU);
                auto *k = smax.kernel("s");
                k->set_mat_scs(true);
                k->register_A(static_cast<unsigned long long>(cs.first),
                              static_cast<unsigned long long>(cs.second), n_rows, n_rows_padded,
                              n_cols, n_chunks, n_elements, nnz, chunk_ptr, chunk_lengths, col,
                              val, perm);
                k->register_B(n_cols, x.data());
                k->register_C(n_rows_padded, y[p].data());
                check(k->run() == 0, "sell run");
            }
            // bit for bit, except that the host's and the GPU's default NaNs
            // differ; finalize copies back the first n_rows entries of the SELL-
            // ordered y, as the CUDA platform does (spmv_cuda_core.cuh:190-200),
            // which hold every row here (no row is empty, so padding sorts last)
            bool same = true;
            for (unsigned long long i = 0; i < n_rows; ++i)
                same &= (std::isnan(y[0][i]) && std::isnan(y[1][i])) ||
                        std::memcmp(&y[0][i], &y[1][i], sizeof(double)) == 0;
            for (unsigned long long i = n_rows; i < n_rows_padded; ++i) same &= y[1][i] == -7.0;
            // the same products written in original row order: row r of y is
            // entry perm[r] of the SELL-ordered one, on both platforms
            for (int p = 0; p < 2; ++p) {
                SMAX::Interface smax;
                std::vector<double> x(m.x), z(n_rows, -7.0);
                smax.register_kernel("s", KernelType::SPMV, p ? PlatformType::HIP : PlatformType::CPU);
                auto *k = smax.kernel("s");
                k->set_mat_scs(true);
                check(k->set_scs_original_order(true) == 0, "original order flag");
                k->register_A(static_cast<unsigned long long>(cs.first),
                              static_cast<unsigned long long>(cs.second), n_rows, n_rows_padded,
                              n_cols, n_chunks, n_elements, nnz, chunk_ptr, chunk_lengths, col,
                              val, perm);
                k->register_B(n_cols, x.data());
                k->register_C(n_rows, z.data());
                check(k->run() == 0, "sell run, original order");
                for (unsigned long long r = 0; r < n_rows; ++r)
                    same &= (std::isnan(z[r]) && std::isnan(y[0][perm[r]])) ||
                            std::memcmp(&z[r], &y[0][perm[r]], sizeof(double)) == 0;
            }
            check(same, "SELL C=" + std::to_string(cs.first) + " sigma=" + std::to_string(cs.second));
            delete[] chunk_ptr;
            delete[] chunk_lengths;
            delete[] col;
            delete[] val;
            delete[] perm;
        }
    }
}

int main(int argc, char **argv) {
    const std::string group = argc > 1 ? argv[1] : "host";
    int failed = 0, passed = 0;
    for (auto &[name, fn] : registry(group)) {
        try {
            fn();
            ++passed;
            std::printf("PASSED %s\n", name.c_str());
        } catch (const std::exception &e) {
            ++failed;
            std::printf("FAILED %s: %s\n", name.c_str(), e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%s: %d passed, %d failed\n", group.c_str(), passed, failed);
    return failed ? 1 : (passed ? 0 : 2);
}
