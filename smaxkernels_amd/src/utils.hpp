// SMAX::Utils: format conversions and timer printing
// (src/utils.hpp:12-94).  convert_crs_to_scs and convert_crs_to_bcrs keep the
// reference signatures and output conventions (arrays allocated with new[]),
// but run on the GPU through libsmax_hip.so; the SELL conversion is
// bit-identical to the reference's, tie order of std::sort included.
#pragma once

#include <cstdint>
#include <cstring>
#include <functional>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>

#include "kernel.hpp"
#include "platforms/hip/hip.hpp"

namespace SMAX {

class Utils {
    std::unordered_map<std::string, std::unique_ptr<Kernel>> &kernels;
    PLATFORM::HIP::Stream stream;

    smax_hip_ctx ctx() {
        if (stream.open()) ErrorHandler::fatal("HIP device unavailable for conversion");
        return stream.get();
    }

    template <typename T> static T *download_new(smax_hip_ctx c, const void *d, size_t n) {
        T *h = new T[n ? n : 1];
        if (n && smax_hip_memcpy_d2h(c, h, d, n * sizeof(T)) != SMAX_HIP_SUCCESS) {
            delete[] h;
            ErrorHandler::fatal(std::string("copy to host: ") + smax_hip_last_error(c));
        }
        return h;
    }

  public:
    UtilitiesContainer *uc = nullptr;
    Utils(std::unordered_map<std::string, std::unique_ptr<Kernel>> &k,
          UtilitiesContainer *u)
        : kernels(k), uc(u) {}

    void print_timers();

    // src/utils/convert_formats.hpp:16-132
    template <typename IT, typename VT, typename ST>
    int convert_crs_to_scs(const ST _n_rows, const ST _n_cols, const ST _nnz,
                           const IT *_col, const IT *_row_ptr, const VT *_val,
                           const ST C, const ST sigma, ST &n_rows, ST &n_rows_padded,
                           ST &n_cols, ST &n_chunks, ST &n_elements, ST &nnz,
                           IT *&chunk_ptr, IT *&chunk_lengths, IT *&col, VT *&val,
                           IT *&perm) {
        static_assert(sizeof(IT) == 4 && std::is_integral_v<IT>,
                      "the HIP conversion stores 32-bit indices");
        static_assert(sizeof(VT) == 2 || sizeof(VT) == 4 || sizeof(VT) == 8,
                      "values must be 2, 4 or 8 bytes");
        smax_hip_ctx c = ctx();
        PLATFORM::HIP::DeviceBuffer drp, dcol, dval;
        if (drp.upload(c, _row_ptr, sizeof(IT) * (_n_rows + 1)) ||
            dcol.upload(c, _col, sizeof(IT) * _nnz) ||
            dval.upload(c, _val, sizeof(VT) * _nnz))
            return 1;
        smax_hip_scs h = nullptr;
        int st = smax_hip_scs_from_crs(c, _n_rows, _n_cols, _nnz, drp.as<int32_t>(),
                                       dcol.as<int32_t>(), dval.get(), sizeof(VT), C,
                                       sigma, &h);
        if (st != SMAX_HIP_SUCCESS)
            ErrorHandler::fatal(std::string("convert_crs_to_scs: ") + smax_hip_last_error(c));
        smax_hip_scs_info in{};
        smax_hip_scs_get_info(h, &in);
        n_rows = static_cast<ST>(in.n_rows);
        n_rows_padded = static_cast<ST>(in.n_rows_padded);
        n_cols = static_cast<ST>(in.n_cols);
        n_chunks = static_cast<ST>(in.n_chunks);
        n_elements = static_cast<ST>(in.n_elements);
        nnz = static_cast<ST>(in.nnz);
        chunk_ptr = download_new<IT>(c, in.chunk_ptr, in.n_chunks + 1);
        chunk_lengths = download_new<IT>(c, in.chunk_lengths, in.n_chunks);
        col = download_new<IT>(c, in.col, in.n_elements);
        val = download_new<VT>(c, in.val, in.n_elements);
        perm = download_new<IT>(c, in.perm, in.n_rows);
        smax_hip_scs_destroy(c, h);
        return 0;
    }

    // src/utils/convert_formats.hpp:134-247 (throws on the same conditions)
    template <typename IT, typename VT, typename ST>
    int convert_crs_to_bcrs(const ST _n_rows, const ST _n_cols, const ST _n_blocks,
                            const IT *_col, const IT *_row_ptr, const VT *_val,
                            ST &n_rows, ST &n_cols, ST &n_blocks, ST &b_height,
                            ST &b_width, ST &height_pad, ST &width_pad, IT *&col,
                            IT *&row_ptr, VT *&val, const ST target_b_height,
                            const ST target_b_width, const ST target_height_pad,
                            const ST target_width_pad, const bool block_column_major) {
        static_assert(sizeof(IT) == 4 && std::is_integral_v<IT>,
                      "the HIP conversion stores 32-bit indices");
        smax_hip_ctx c = ctx();
        const ST nnz = _n_blocks;  // the CRS input's nonzero count
        PLATFORM::HIP::DeviceBuffer drp, dcol, dval;
        if (drp.upload(c, _row_ptr, sizeof(IT) * (_n_rows + 1)) ||
            dcol.upload(c, _col, sizeof(IT) * nnz) ||
            dval.upload(c, _val, sizeof(VT) * nnz))
            return 1;
        smax_hip_bcrs h = nullptr;
        int st = smax_hip_bcrs_from_crs(c, _n_rows, _n_cols, nnz, drp.as<int32_t>(),
                                        dcol.as<int32_t>(), dval.get(), sizeof(VT),
                                        target_b_height, target_b_width,
                                        target_height_pad, target_width_pad,
                                        block_column_major ? 1 : 0, &h);
        if (st == SMAX_HIP_ERROR_INVALID_VALUE)
            throw std::runtime_error(smax_hip_last_error(c));
        if (st != SMAX_HIP_SUCCESS)
            ErrorHandler::fatal(std::string("convert_crs_to_bcrs: ") + smax_hip_last_error(c));
        smax_hip_bcrs_info in{};
        smax_hip_bcrs_get_info(h, &in);
        n_rows = static_cast<ST>(in.n_block_rows);
        n_cols = static_cast<ST>(in.n_block_cols);
        n_blocks = static_cast<ST>(in.n_blocks);
        b_height = static_cast<ST>(in.b_h);
        b_width = static_cast<ST>(in.b_w);
        height_pad = static_cast<ST>(in.h_pad);
        width_pad = static_cast<ST>(in.w_pad);
        row_ptr = download_new<IT>(c, in.row_ptr, in.n_block_rows + 1);
        col = download_new<IT>(c, in.col, in.n_blocks);
        val = download_new<VT>(c, in.val, in.n_blocks * in.h_pad * in.w_pad);
        smax_hip_bcrs_destroy(c, h);
        return 0;
    }
};

} // namespace SMAX
