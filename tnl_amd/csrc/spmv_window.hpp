// spmv_window.hpp -- the windowed SpMV plan (spmv_window.hip).
#pragma once

#include "common.hpp"

namespace tnlmi {

struct WindowPlan
{
   int64_t rows = 0, columns = 0, nnz = 0;
   tnlmi_dtype value_type = TNLMI_F64;
   int32_t win = 8192;  // columns per window (x entries staged in LDS by phase 1)
   int32_t segs = 512;  // phase 2: windows per sweep of a row block (256 or 512)
   int32_t per = 4;     // phase 2: consecutive terms per lane and load (4: 8-lane groups; 2: 16-lane)
   int32_t unroll = 4;  // phase 1: nonzeros per thread and batch
   bool align = true;    // phase 2 reads products and slots from aligned positions
   bool across = true;   // phase 2: a wave's groups read adjacent windows' chunks in each round
   int32_t K = 0;    // column windows
   int64_t B = 0;     // row blocks
   int64_t H = 0;     // rows longer than cap (cut into pieces)
   int64_t cap = 0;   // nonzeros per row block at most
   DeviceArray< int64_t > wstart;  // K + 2: first sorted nonzero of each window (window K: padding slots)
   DeviceArray< int32_t > seg;     // (B + 1) x K: first nonzero of row block b in window k
   DeviceArray< int32_t > offs;    // rows + 1: the CSR offsets
   DeviceArray< uint16_t > rorder; // rows: each run's rows by length, longest first (index inside the run)
   DeviceArray< int32_t > hdr;     // 8 B: rows, nonzeros, piece and long-row range of each row block
   DeviceArray< int32_t > lrows;   // 3 per row longer than the one-thread limit: row, first and end slot
   DeviceArray< int32_t > huge;    // 3 H: (row, first piece block, pieces) of each row longer than cap
   DeviceArray< char > part;       // B: partial sums of the piece blocks (compute type)
   DeviceArray< char > pv;         // nonzeros sorted by window (rows ascending inside), value type
   DeviceArray< uint16_t > pc;     // their column inside the window
   DeviceArray< uint16_t > dest;   // CSR slot inside its row block of each product, block-major: a row
                                   // block's chunks (one per window) side by side
   DeviceArray< int32_t > dbase;   // B x K: slot of chunk (b, k)'s first product minus its index in prod
   DeviceArray< char > prod;       // phase-1 products (compute type)
};

void window_build( WindowPlan& W, const tnlmi_matrix& A, hipStream_t s );
void window_execute( const WindowPlan& W, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct, double alpha,
                     double beta, int64_t begin, int64_t end, hipStream_t s );

}  // namespace tnlmi
