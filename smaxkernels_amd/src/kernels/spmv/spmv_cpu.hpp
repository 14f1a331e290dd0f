// CPU platform of the SpMV and SpMM kernels: the reference's OpenMP kernels
// (src/kernels/spmv/spmv_cpu/spmv_cpu_{crs,scs,bcrs}_impl.hpp,
// src/kernels/spmm/spmm_cpu/spmm_cpu_crs_impl.hpp) and their type dispatch
// over every IntType x {FLOAT32, FLOAT64} (src/kernels/spmv/spmv_cpu.hpp:38-96),
// plus BFLOAT16 values with FLOAT32 vectors as on the HIP platform.
//
// Every output is one sequential fma chain over its row in stored order --
// the reference's row sum as its Release build (-march=native) contracts it --
// so the CPU and HIP platforms give the same results bit for bit.  It is a
// platform chosen by PlatformType::CPU, the reference's default; it never
// stands in for the HIP platform.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../common.hpp"
#include "spmv_common.hpp"

namespace SMAX::KERNELS::CPU {

inline float widen(uint16_t bf16) {
    const uint32_t u = static_cast<uint32_t>(bf16) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline float widen(float v) { return v; }
inline double widen(double v) { return v; }

inline float fma_(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }

// naive_crs_spmv (spmv_cpu_crs_impl.hpp:9-38)
template <typename IT, typename V, typename A>
void crs_spmv(ULL n_rows, const IT *col, const IT *row_ptr, const V *val, const A *x, A *y) {
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < static_cast<long long>(n_rows); ++r) {
        A acc = 0;
        for (ULL j = static_cast<ULL>(row_ptr[r]); j < static_cast<ULL>(row_ptr[r + 1]); ++j)
            acc = fma_(widen(val[j]), x[col[j]], acc);
        y[r] = acc;
    }
}

// naive_scs_spmv (spmv_cpu_scs_impl.hpp:9-49): y in SELL order, every slot
// chained over its chunk's full length (padding adds 0 * x[0]).
template <typename IT, typename V, typename A>
void scs_spmv(ULL C, ULL n_chunks, const IT *chunk_ptr, const IT *chunk_lengths, const IT *col,
              const V *val, const A *x, A *y) {
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < static_cast<long long>(n_chunks); ++c) {
        const ULL base = static_cast<ULL>(chunk_ptr[c]);
        const ULL len = static_cast<ULL>(chunk_lengths[c]);
        for (ULL i = 0; i < C; ++i) {
            A acc = 0;
            for (ULL j = 0; j < len; ++j) {
                const ULL o = base + j * C + i;
                acc = fma_(widen(val[o]), x[col[o]], acc);
            }
            y[c * C + i] = acc;
        }
    }
}

// naive_bcrs_spmv (spmv_cpu_bcrs_impl.hpp:19-85): per block row, output ix
// chains over the row's blocks in order and each block's columns jx in
// ascending order; column-major blocks store (ix, jx) at jx*h_pad + ix.
template <typename IT, typename V, typename A>
void bcrs_spmv(const BCRSMatrix &m, bool col_major, const IT *col, const IT *row_ptr,
               const V *val, const A *x, A *y) {
    const ULL bh = m.b_height, bw = m.b_width, hp = m.height_pad, wp = m.width_pad;
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < static_cast<long long>(m.n_rows); ++r) {
        for (ULL ix = 0; ix < bh; ++ix) {
            A acc = 0;
            for (ULL j = static_cast<ULL>(row_ptr[r]); j < static_cast<ULL>(row_ptr[r + 1]); ++j)
                for (ULL jx = 0; jx < bw; ++jx) {
                    const ULL o = j * hp * wp + (col_major ? jx * hp + ix : ix * wp + jx);
                    acc = fma_(widen(val[o]), x[static_cast<ULL>(col[j]) * wp + jx], acc);
                }
            y[r * hp + ix] = acc;
        }
    }
}

// naive_crs_spmm_{col,row}_maj (spmm_cpu_crs_impl.hpp:9-89): per row a
// tmp[k], then for each stored entry a pass over the k columns; column-major
// operands use n_rows as leading dimension, as the reference does.
template <typename IT, typename V, typename A>
void crs_spmm(ULL n_rows, ULL k, bool row_major, const IT *col, const IT *row_ptr, const V *val,
              const A *X, A *Y) {
#pragma omp parallel
    {
        std::vector<A> tmp(k);
#pragma omp for schedule(static)
        for (long long r = 0; r < static_cast<long long>(n_rows); ++r) {
            for (ULL v = 0; v < k; ++v) tmp[v] = 0;
            for (ULL j = static_cast<ULL>(row_ptr[r]); j < static_cast<ULL>(row_ptr[r + 1]); ++j) {
                const A a = widen(val[j]);
                const ULL c = static_cast<ULL>(col[j]);
                for (ULL v = 0; v < k; ++v)
                    tmp[v] = fma_(a, X[row_major ? c * k + v : n_rows * v + c], tmp[v]);
            }
            for (ULL v = 0; v < k; ++v) {
                if (row_major)
                    Y[r * k + v] = tmp[v];
                else
                    Y[r + v * n_rows] = tmp[v];
            }
        }
    }
}

// Calls f.template run<IT, V, A>() for the registered types (the reference's
// dispatch_cpu, spmv_cpu.hpp:38-96).  Returns 1 for a type pair that has no
// kernel.
template <typename IT, typename F> int dispatch_value(FloatType t, F &f) {
    switch (t) {
    case FloatType::FLOAT64: f.template run<IT, double, double>(); return 0;
    case FloatType::FLOAT32: f.template run<IT, float, float>(); return 0;
    case FloatType::BFLOAT16: f.template run<IT, uint16_t, float>(); return 0;
    }
    return 1;
}

template <typename F> int dispatch(const KernelContext &k, F &f) {
    switch (k.int_type) {
    case IntType::INT16: return dispatch_value<int16_t>(k.float_type, f);
    case IntType::INT32: return dispatch_value<int32_t>(k.float_type, f);
    case IntType::INT64: return dispatch_value<int64_t>(k.float_type, f);
    case IntType::UINT16: return dispatch_value<uint16_t>(k.float_type, f);
    case IntType::UINT32: return dispatch_value<uint32_t>(k.float_type, f);
    case IntType::UINT64: return dispatch_value<uint64_t>(k.float_type, f);
    }
    return 1;
}

// The SpMV apply of the CPU platform on the registered host arrays; x and y
// offsets are element offsets, as on the other platforms
// (spmv_cpu_core.hpp:62-63).
struct SpMVApply {
    const SPMV::Args &a;
    const SPMV::Flags &fl;
    ULL x_off, y_off;
    template <typename IT, typename V, typename A> void run() const {
        const A *x = static_cast<const A *>(a.x.val) + x_off;
        A *y = static_cast<A *>(a.y.val) + y_off;
        if (fl.is_mat_scs) {
            const SCSMatrix &m = *a.A.scs;
            std::vector<A> tmp;
            A *ys = y;
            if (fl.scs_original_order) {
                tmp.resize(m.n_chunks * m.C);
                ys = tmp.data();
            }
            scs_spmv<IT, V, A>(m.C, m.n_chunks, static_cast<const IT *>(m.chunk_ptr),
                               static_cast<const IT *>(m.chunk_lengths),
                               static_cast<const IT *>(m.col), static_cast<const V *>(m.val), x,
                               ys);
            if (fl.scs_original_order) {
                // row r sits at slot perm[r] (convert_formats.hpp:81-86); no perm: r
                const IT *perm = static_cast<const IT *>(m.perm);
                for (ULL r = 0; r < m.n_rows; ++r)
                    y[r] = tmp[perm ? static_cast<ULL>(perm[r]) : r];
            }
        } else if (fl.is_mat_bcrs) {
            const BCRSMatrix &m = *a.A.bcrs;
            bcrs_spmv<IT, V, A>(m, fl.is_block_column_major, static_cast<const IT *>(m.col),
                                static_cast<const IT *>(m.row_ptr),
                                static_cast<const V *>(m.val), x, y);
        } else {
            const CRSMatrix &m = *a.A.crs;
            crs_spmv<IT, V, A>(m.n_rows, static_cast<const IT *>(m.col),
                               static_cast<const IT *>(m.row_ptr), static_cast<const V *>(m.val),
                               x, y);
        }
    }
};

inline int spmv_apply(const KernelContext &k, const SPMV::Args &a, const SPMV::Flags &fl,
                      ULL x_off, ULL y_off) {
    SpMVApply f{a, fl, x_off, y_off};
    return dispatch(k, f);
}

// The SpMM apply: Y = A X with X/Y at element offsets, k = X's column count.
struct SpMMApply {
    const CRSMatrix &m;
    const void *X;
    void *Y;
    ULL k;
    bool row_major;
    template <typename IT, typename V, typename A> void run() const {
        crs_spmm<IT, V, A>(m.n_rows, k, row_major, static_cast<const IT *>(m.col),
                           static_cast<const IT *>(m.row_ptr), static_cast<const V *>(m.val),
                           static_cast<const A *>(X), static_cast<A *>(Y));
    }
};

} // namespace SMAX::KERNELS::CPU
