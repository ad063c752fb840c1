// spmv_tile.hpp -- the wavefront tile of the element-split CSR kernels
// (shared by the plain CSR path and the XCD column-split plan).
#pragma once

#include "common.hpp"
#include "segments.hpp"

namespace tnlmi {

// LDS ordering between lanes of one wavefront (no workgroup barrier needed).
__device__ __forceinline__ void wave_sync()
{
   __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

// Result store of one row (Matrices/SparseMatrixBase.hpp:346-389).
template< typename X, typename C >
__device__ __forceinline__ void keep( X* __restrict__ y, int64_t r, C v, C alpha, C beta )
{
   if( beta == C( 0 ) )
      y[ r ] = alpha == C( 1 ) ? X( v ) : X( alpha * v );
   else
      y[ r ] = alpha == C( 1 ) ? X( beta * (C) y[ r ] + v ) : X( beta * (C) y[ r ] + alpha * v );
}


// Tile t of a CSR element stream (offsets `off`, n_rows rows, nnz elements):
// the elements [t*TILE, (t+1)*TILE) and the rows [rs, re) whose first
// element lies there.
// Phase 1 streams the tile's columns/values once with non-temporal, fully
// coalesced loads (TILE/64 independent loads per lane in flight), gathers
// x and stages the products in LDS (s, TILE entries of this wavefront).
// Phase 2: carry(rc, v) receives the partial sum of the row rc that started
// in an earlier tile (rc = -1 if none) -- called by lane 0 only; store(r, v)
// the sum of each owned row's part in this tile -- 64/pow2(rows) lanes per
// row, or one lane per row summing left to right from 0 (the reference's host
// order) when the tile holds 64 rows or more.
// WHOLE (every row at most kWholeRowMax elements long): the owner of the row
// that continues past the tile end also loads that row's remaining elements
// (at most tail_max = longest row - 1, one per lane, beside the tile's own
// loads) and stages their products after the tile's in s, so every owned row
// is stored complete and carry is never called; every row is then summed by
// one lane, left to right from 0.  s then holds TILE + kWholeRowMax entries.
template< typename V, typename X, typename C, bool WHOLE = false, int TILE = kTileNnz, typename Store,
          typename Carry >
__device__ __forceinline__ void csr_tile( const int32_t* __restrict__ off, int64_t n_rows,
                                          const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                          const X* __restrict__ x, int64_t t, int64_t rs, int64_t re, int64_t nnz,
                                          C* __restrict__ s, int lane, Store&& store, Carry&& carry,
                                          int32_t tail_max = 0 )
{
   constexpr int kTileUnroll = TILE / 64;
   const int64_t e0 = t * TILE;
   const int64_t e1 = min( e0 + TILE, nnz );
   const int32_t nel = (int32_t) max( (int64_t) 0, e1 - e0 );
   // WHOLE: elements [e1, off[re]) finish the last owned row; off[re] - e1 < the
   // row's length <= tail_max + 1 because that row starts before e1
   int32_t ct = -1;
   C vt = 0;
   if( WHOLE && lane < tail_max && e1 + lane < nnz ) {
      ct = cols[ e1 + lane ];
      vt = load_val< V, C >( vals, e1 + lane );
   }
   // row bounds of this lane's first row in phase 2, loaded beside the tile
   // (phase 2 then waits on no further memory for rows of >= 64-row tiles'
   // first pass, nor for any row of a smaller tile)
   const int32_t nrows = (int32_t) ( re - rs );
   int G = 64;  // lanes per row: largest power of two with nrows * G <= 64 (1 for >= 64 rows or WHOLE)
   while( G > 1 && ( WHOLE || nrows * G > 64 ) )
      G >>= 1;
   const int64_t r_first = rs + lane / G;
   int64_t fb = 0, fe = 0;
   if( r_first < re ) {
      fb = off[ r_first ];
      fe = off[ r_first + 1 ];
   }
   const int64_t tail = WHOLE && re > rs ? (int64_t) off[ re ] - e1 : 0;
   {
      const int32_t* __restrict__ cp = cols + e0;
      const V* __restrict__ vp = vals + e0;
      int32_t c[ kTileUnroll ];
      C v[ kTileUnroll ];
#pragma unroll
      for( int u = 0; u < kTileUnroll; u++ ) {
         const int32_t i = lane + 64 * u;
         c[ u ] = i < nel ? nt_load( cp + i ) : -1;
      }
#pragma unroll
      for( int u = 0; u < kTileUnroll; u++ ) {
         const int32_t i = lane + 64 * u;
         v[ u ] = i < nel ? nt_load_val< V, C >( vp, i ) : C( 0 );
      }
#pragma unroll
      for( int u = 0; u < kTileUnroll; u++ ) {
         const int32_t i = lane + 64 * u;
         if( i < nel )
            s[ i ] = c[ u ] >= 0 ? v[ u ] * (C) x[ c[ u ] ] : C( 0 );
      }
      if( WHOLE && lane < tail )
         s[ nel + lane ] = ct >= 0 ? vt * (C) x[ ct ] : C( 0 );
   }
   wave_sync();
   if( ! WHOLE ) {
      // the row containing element e0 started earlier iff no owned row starts at e0
      int64_t rc = -1;
      int32_t cend = 0;
      if( e0 < nnz && rs > 0 && ( rs == n_rows || off[ rs ] > e0 ) ) {
         rc = rs - 1;
         cend = (int32_t) ( min( (int64_t) off[ rs ], e1 ) - e0 );
      }
      C acc = 0;
      if( rc >= 0 )
         for( int32_t i = lane; i < cend; i += 64 )
            acc += s[ i ];
      acc = group_sum< 64 >( acc );
      if( lane == 0 )
         carry( rc, acc );
   }
   if( WHOLE || nrows >= 64 ) {
      for( int64_t r = rs + lane; r < re; r += 64 ) {
         if( r != r_first ) {
            fb = off[ r ];
            fe = off[ r + 1 ];
         }
         const int32_t b = (int32_t) ( fb - e0 ), e = (int32_t) ( ( WHOLE ? fe : min( fe, e1 ) ) - e0 );
         C acc = 0;
         for( int32_t i = b; i < e; i++ )
            acc = acc + s[ i ];
         store( r, acc );
      }
   }
   else if( nrows > 0 ) {
      const int grp = lane / G, gl = lane % G;
      C acc = 0;
      const int64_t r = r_first;
      if( grp < nrows ) {
         const int32_t b = (int32_t) ( fb - e0 ), e = (int32_t) ( ( WHOLE ? fe : min( fe, e1 ) ) - e0 );
         for( int32_t i = b + gl; i < e; i += G )
            acc += s[ i ];
      }
      acc = group_sum_dyn( acc, G );
      if( gl == 0 && grp < nrows )
         store( r, acc );
   }
}

}  // namespace tnlmi
