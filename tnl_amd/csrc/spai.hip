// spai.hip -- the learned sparse approximate inverse: a message-passing
// network over the graph of A (an edge per nonzero, CSR order) whose output
// is M = D^-1 (I + O) on A's own pattern, a left preconditioner for GMRES.
//
// Network (hidden 64, 3 layers), every 64 x 64 dense layer a BF16 GEMM on
// the matrix cores (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation), BF16
// activations:
//   encoder  E = relu(x_e Ae + ae), H = relu(x_i An + an)      (4 inputs, FP32)
//   layer    P = H Wu, Q = H Wv                                 (FP32)
//            E = relu(E We + P[col] + Q[row] + be)              (fused epilogue)
//            G = mean of the row's E
//            H = relu(H Wh + G Wg + bh)                         (two-term GEMM)
//   decoder  o = E wo + bo,  M_ij = (delta_ij + o_ij) / a_ii
// Training: probe loss L = sum_k ||z_k - M A z_k||^2 / (K n) over Rademacher
// z_k, hand-written backward kernels (MFMA GEMMs with transposed weights for
// the input gradients, LDS-staged MFMA X^T dY for the weight gradients with
// per-block partial sums reduced in block order: deterministic), Adam.
//
// No reference counterpart (TNL has no learned preconditioner).  The CPU
// specification is oracle/spai_oracle.py.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "segments.hpp"
#include "spmv.hpp"

namespace tnlmi {
namespace {

constexpr int kH = 64;       // hidden width
constexpr int kL = 3;        // message-passing layers
constexpr int kMats = 5;     // We, Wu, Wv, Wh, Wg per layer
constexpr int kWRows = 256;  // rows per block of the weight-gradient kernel

// flat parameter layout (the oracle's param_shapes order)
constexpr int64_t oAe = 0, oae = 256, oAn = 320, oan = 576, oLayer0 = 640;
constexpr int64_t kLayerSize = kMats * kH * kH + 2 * kH;
constexpr int64_t owo = oLayer0 + kL * kLayerSize, obo = owo + kH;
constexpr int64_t kParams = obo + 1;
__host__ __device__ constexpr int64_t oW( int l, int m )
{
   return oLayer0 + l * kLayerSize + (int64_t) m * kH * kH;
}
__host__ __device__ constexpr int64_t obe( int l )
{
   return oLayer0 + l * kLayerSize + kMats * kH * kH;
}
__host__ __device__ constexpr int64_t obh( int l )
{
   return obe( l ) + kH;
}
enum { mWe = 0, mWu = 1, mWv = 2, mWh = 3, mWg = 4 };

typedef short s4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef float f4 __attribute__( ( ext_vector_type( 4 ) ) );

__device__ __forceinline__ uint16_t bf( float f )
{
   return float_to_bf16( f );
}
__device__ __forceinline__ float fb( uint16_t b )
{
   return bf16_to_float( b );
}

// ------------------------------------------------------------- dense layers
// Y[R x 64] = sum_t X_t[R x 64] B_t[64 x 64], X_t BF16 row-major, B_t given
// as Bn[n][k] = B[k][n] (BF16).  Block: 4 wavefronts x 16 rows; each
// wavefront the full 64 columns (4 column blocks x 4 k-steps of 16).
// MFMA 16x16x16 operand layout: A lane l holds A[l % 16][4 (l / 16) + 0..3],
// B lane l holds B[4 (l / 16) + 0..3][l % 16], C lane l holds
// C[4 (l / 16) + i][l % 16], i = 0..3.
struct Term
{
   const uint16_t* X;
   const uint16_t* Bn;
};
enum Epi { EPI_STORE = 0, EPI_ADD = 1, EPI_EDGE = 2, EPI_NODE = 3 };
struct EpiArgs
{
   float* yf;            // EPI_STORE / EPI_ADD (FP32 out)
   uint16_t* yb;         // EPI_EDGE / EPI_NODE (BF16 out)
   const float* bias;    // EPI_EDGE / EPI_NODE
   const float* P;       // EPI_EDGE: + P[col[r]] + Q[row[r]]
   const float* Q;
   const int32_t* col;
   const int32_t* row;
};

template< int NT, int EPI >
__global__ void __launch_bounds__( 256 ) gemm64( int64_t R, Term t0, Term t1, Term t2, EpiArgs e )
{
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int64_t r0 = (int64_t) blockIdx.x * 64 + w * 16;
   const int lr = lane & 15, lk = 4 * ( lane >> 4 );
   f4 acc[ 4 ];
#pragma unroll
   for( int cb = 0; cb < 4; cb++ )
      acc[ cb ] = f4{ 0.f, 0.f, 0.f, 0.f };
   const Term ts[ 3 ] = { t0, t1, t2 };
#pragma unroll
   for( int t = 0; t < NT; t++ ) {
      const int64_t ra = r0 + lr;
#pragma unroll
      for( int kb = 0; kb < 4; kb++ ) {
         s4 a = s4{ 0, 0, 0, 0 };
         if( ra < R )
            a = *(const s4*) ( ts[ t ].X + ra * kH + kb * 16 + lk );
#pragma unroll
         for( int cb = 0; cb < 4; cb++ ) {
            const s4 b = *(const s4*) ( ts[ t ].Bn + ( cb * 16 + lr ) * kH + kb * 16 + lk );
            acc[ cb ] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k( a, b, acc[ cb ], 0, 0, 0 );
         }
      }
   }
#pragma unroll
   for( int cb = 0; cb < 4; cb++ ) {
      const int c = cb * 16 + lr;
#pragma unroll
      for( int i = 0; i < 4; i++ ) {
         const int64_t r = r0 + lk + i;
         if( r >= R )
            continue;
         const float v = acc[ cb ][ i ];
         if( EPI == EPI_STORE )
            e.yf[ r * kH + c ] = v;
         else if( EPI == EPI_ADD )
            e.yf[ r * kH + c ] += v;
         else if( EPI == EPI_EDGE ) {
            const int32_t j = e.col[ r ];  // -1: a padding slot, no edge
            const float pj = j >= 0 ? e.P[ (int64_t) j * kH + c ] : 0.f;
            e.yb[ r * kH + c ] = bf( fmaxf( v + pj + e.Q[ (int64_t) e.row[ r ] * kH + c ] + e.bias[ c ], 0.f ) );
         }
         else
            e.yb[ r * kH + c ] = bf( fmaxf( v + e.bias[ c ], 0.f ) );
      }
   }
}

// dW[64 x 64] = X^T dY over R rows (and the column sums of dY): per block of
// kWRows rows a partial product from LDS-staged tiles, summed over the blocks
// in a fixed order by reduce_partials (chunks of consecutive blocks, then the
// chunk sums).
__global__ void __launch_bounds__( 256 ) wgrad64( int64_t R, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ dY, float* __restrict__ part,
                                                  float* __restrict__ bpart )
{
   __shared__ uint16_t sx[ kWRows * kH ], sy[ kWRows * kH ];
   const int64_t base = (int64_t) blockIdx.x * kWRows;
   // stage the tiles (16 bytes per thread per step), zero past R
   for( int q = threadIdx.x; q < kWRows * kH / 8; q += 256 ) {
      const int64_t r = base + q / 8;
      typedef int i4 __attribute__( ( ext_vector_type( 4 ) ) );
      i4 vx = i4{ 0, 0, 0, 0 }, vy = i4{ 0, 0, 0, 0 };
      if( r < R ) {
         vx = *(const i4*) ( X + base * kH + (int64_t) q * 8 );
         vy = *(const i4*) ( dY + base * kH + (int64_t) q * 8 );
      }
      *(i4*) ( sx + q * 8 ) = vx;
      *(i4*) ( sy + q * 8 ) = vy;
   }
   __syncthreads();
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int lr = lane & 15, lk = 4 * ( lane >> 4 );
   f4 acc[ 4 ];
#pragma unroll
   for( int cb = 0; cb < 4; cb++ )
      acc[ cb ] = f4{ 0.f, 0.f, 0.f, 0.f };
   const int m = w * 16 + lr;  // feature of X = row of dW
   for( int ks = 0; ks < kWRows / 16; ks++ ) {
      const int k = ks * 16 + lk;  // row of the tile
      const s4 a = s4{ (short) sx[ ( k + 0 ) * kH + m ], (short) sx[ ( k + 1 ) * kH + m ],
                       (short) sx[ ( k + 2 ) * kH + m ], (short) sx[ ( k + 3 ) * kH + m ] };
#pragma unroll
      for( int cb = 0; cb < 4; cb++ ) {
         const int n = cb * 16 + lr;
         const s4 b = s4{ (short) sy[ ( k + 0 ) * kH + n ], (short) sy[ ( k + 1 ) * kH + n ],
                          (short) sy[ ( k + 2 ) * kH + n ], (short) sy[ ( k + 3 ) * kH + n ] };
         acc[ cb ] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k( a, b, acc[ cb ], 0, 0, 0 );
      }
   }
   float* out = part + (int64_t) blockIdx.x * kH * kH;
#pragma unroll
   for( int cb = 0; cb < 4; cb++ )
#pragma unroll
      for( int i = 0; i < 4; i++ )
         out[ ( w * 16 + lk + i ) * kH + cb * 16 + lr ] = acc[ cb ][ i ];
   if( bpart && threadIdx.x < kH ) {
      float sum = 0.f;
      for( int r = 0; r < kWRows; r++ )
         sum += fb( sy[ r * kH + threadIdx.x ] );
      bpart[ (int64_t) blockIdx.x * kH + threadIdx.x ] = sum;
   }
}

// out[e] = sum over blocks b (in order) of part[b * width + e]
__global__ void reduce_parts( const float* __restrict__ part, int64_t nblk, int width, float* __restrict__ out )
{
   const int e = blockIdx.x * blockDim.x + threadIdx.x;
   if( e >= width )
      return;
   float s = 0.f;
   for( int64_t b = 0; b < nblk; b++ )
      s += part[ b * width + e ];
   out[ e ] = s;
}

// Two-level form of reduce_parts for long partial lists: chunk c of at most
// L consecutive blocks per column (in block order) by one lane each, then the
// chunk sums in chunk order -- a fixed order, so deterministic, with
// (columns / 64) x chunks waves in the first level instead of one thread
// walking every block.
constexpr int kRedChunks = 64;

__global__ void __launch_bounds__( 64 ) reduce_parts_chunks( const float* __restrict__ part, int64_t nblk, int width,
                                                             int64_t L, float* __restrict__ out )
{
   const int e = blockIdx.x * 64 + threadIdx.x;
   if( e >= width )
      return;
   const int64_t c = blockIdx.y;
   const int64_t b0 = c * L, b1 = std::min< int64_t >( nblk, b0 + L );
   float s = 0.f;
   for( int64_t b = b0; b < b1; b++ )
      s += part[ b * width + e ];
   out[ c * width + e ] = s;
}

// ------------------------------------------------------------ graph and IO
struct Graph
{
   int64_t n = 0, E = 0;
   DeviceArray< int32_t > erow, cperm, cptr, deg;
   DeviceArray< double > diag;
   DeviceArray< float > xe, xn;
   DeviceArray< uint8_t > isd;
};

__global__ void g_rows( int64_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ cols,
                        const double* __restrict__ vals, int32_t* __restrict__ erow, int32_t* __restrict__ deg,
                        double* __restrict__ diag, uint8_t* __restrict__ isd )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= n )
      return;
   double d = 0.0;
   for( int32_t e = off[ i ]; e < off[ i + 1 ]; e++ ) {
      erow[ e ] = (int32_t) i;
      const bool dg = cols[ e ] == i;
      isd[ e ] = dg ? 1 : 0;
      if( dg )
         d = vals[ e ];
   }
   deg[ i ] = off[ i + 1 ] - off[ i ];
   diag[ i ] = d;
}

__global__ void g_features( int64_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ cols,
                            const double* __restrict__ vals, const double* __restrict__ diag,
                            float* __restrict__ xe, float* __restrict__ xn )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= n )
      return;
   const double di = diag[ i ];
   double oabs = 0.0, rs = 0.0;
   for( int32_t e = off[ i ]; e < off[ i + 1 ]; e++ ) {
      const int32_t j = cols[ e ];
      const double a = vals[ e ];
      if( j < 0 ) {  // padding slot: no edge
         for( int k = 0; k < 4; k++ )
            xe[ (int64_t) e * 4 + k ] = 0.f;
         continue;
      }
      xe[ (int64_t) e * 4 + 0 ] = (float) ( a / di );
      xe[ (int64_t) e * 4 + 1 ] = (float) ( a / diag[ j ] );
      xe[ (int64_t) e * 4 + 2 ] = j == i ? 1.f : 0.f;
      xe[ (int64_t) e * 4 + 3 ] = (float) ( fabs( a ) / di );
      if( j != i )
         oabs += fabs( a );
      rs += a;
   }
   xn[ i * 4 + 0 ] = (float) ( oabs / di );
   xn[ i * 4 + 1 ] = (float) ( rs / di );
   xn[ i * 4 + 2 ] = (float) ( off[ i + 1 ] - off[ i ] ) / 5.f;
   xn[ i * 4 + 3 ] = 1.f;
}

__global__ void g_colptr( int64_t E, int64_t n, const int32_t* __restrict__ skeys, int32_t* __restrict__ cptr )
{
   const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p > E )
      return;
   const int32_t prev = p == 0 ? -1 : skeys[ p - 1 ];
   const int32_t cur = p == E ? (int32_t) n : skeys[ p ];
   for( int32_t c = prev + 1; c <= cur; c++ )
      cptr[ c ] = (int32_t) p;
}

__global__ void iota_k( int64_t E, int32_t* __restrict__ v )
{
   const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p < E )
      v[ p ] = (int32_t) p;
}

// ------------------------------------------------------------ forward parts
// relu(x W + b) for K = 4 inputs (FP32), BF16 out
__global__ void encode_k( int64_t R, const float* __restrict__ x, const float* __restrict__ W,
                          const float* __restrict__ b, uint16_t* __restrict__ y )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= R * kH )
      return;
   const int64_t r = q / kH;
   const int c = (int) ( q % kH );
   float v = b[ c ];
#pragma unroll
   for( int k = 0; k < 4; k++ )
      v += x[ r * 4 + k ] * W[ k * kH + c ];
   y[ q ] = bf( fmaxf( v, 0.f ) );
}

// G = BF16 of the mean of the row's edge features
__global__ void mean_k( int64_t n, const int32_t* __restrict__ off, const uint16_t* __restrict__ E,
                        uint16_t* __restrict__ G )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= n * kH )
      return;
   const int64_t i = q / kH;
   const int c = (int) ( q % kH );
   float s = 0.f;
   for( int32_t e = off[ i ]; e < off[ i + 1 ]; e++ )
      s += fb( E[ (int64_t) e * kH + c ] );
   const int32_t d = off[ i + 1 ] - off[ i ];
   G[ q ] = bf( d > 0 ? s / (float) d : 0.f );
}

// o_e = E_e . wo + bo (FP32), M_e = (delta + o) / a_ii (FP64)
__global__ void decode_k( int64_t E, const uint16_t* __restrict__ X, const float* __restrict__ wo,
                          const float* __restrict__ bo, const int32_t* __restrict__ erow,
                          const int32_t* __restrict__ cols, const uint8_t* __restrict__ isd,
                          const double* __restrict__ diag, double* __restrict__ M )
{
   const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( e >= E )
      return;
   if( cols[ e ] < 0 ) {
      M[ e ] = 0.0;
      return;
   }
   float o = bo[ 0 ];
   for( int c = 0; c < kH; c++ )
      o += fb( X[ e * kH + c ] ) * wo[ c ];
   M[ e ] = ( (double) isd[ e ] + (double) o ) / diag[ erow[ e ] ];
}

// ------------------------------------------------------------ probe loss
__device__ __forceinline__ uint64_t mix64( uint64_t z )
{
   z += 0x9e3779b97f4a7c15ull;
   z = ( z ^ ( z >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
   z = ( z ^ ( z >> 27 ) ) * 0x94d049bb133111ebull;
   return z ^ ( z >> 31 );
}
__global__ void rademacher_k( int64_t cnt, uint64_t seed, double* __restrict__ Z )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q < cnt )
      Z[ q ] = ( mix64( seed * 0x2545f4914f6cdd1dull + (uint64_t) q ) >> 63 ) ? 1.0 : -1.0;
}

// y = V x over the CSR pattern (one thread per row)
__global__ void csr_mv( int64_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ cols,
                        const double* __restrict__ V, const double* __restrict__ x, double* __restrict__ y )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= n )
      return;
   double s = 0.0;
   for( int32_t e = off[ i ]; e < off[ i + 1 ]; e++ )
      if( cols[ e ] >= 0 )
         s += V[ e ] * x[ cols[ e ] ];
   y[ i ] = s;
}

// r = z - w (in place over w), partial sums of r.r per block
__global__ void __launch_bounds__( 256 ) resid_k( int64_t n, const double* __restrict__ z, double* __restrict__ w,
                                                  double* __restrict__ part )
{
   __shared__ double sm[ 4 ];
   const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
   double v = 0.0;
   if( i < n ) {
      const double r = z[ i ] - w[ i ];
      w[ i ] = r;
      v = r * r;
   }
   v = block_sum< 256 >( v, sm );
   if( threadIdx.x == 0 )
      part[ blockIdx.x ] = v;
}

__global__ void sum_k( const double* __restrict__ part, int64_t np, double scale, double* __restrict__ out )
{
   __shared__ double sm[ 4 ];
   double v = 0.0;
   for( int64_t i = threadIdx.x; i < np; i += 256 )
      v += part[ i ];
   v = block_sum< 256 >( v, sm );
   if( threadIdx.x == 0 )
      out[ 0 ] = v * scale;
}

// dL/dM_e = -2 / (K n) sum_k r_k[row] y_k[col]
__global__ void dM_k( int64_t E, int64_t n, int K, const int32_t* __restrict__ erow, const int32_t* __restrict__ cols,
                      const double* __restrict__ Rk, const double* __restrict__ Yk, double* __restrict__ dM )
{
   const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( e >= E )
      return;
   double s = 0.0;
   if( cols[ e ] >= 0 )
      for( int k = 0; k < K; k++ )
         s += Rk[ (int64_t) k * n + erow[ e ] ] * Yk[ (int64_t) k * n + cols[ e ] ];
   dM[ e ] = -2.0 * s / ( (double) K * (double) n );
}

// ------------------------------------------------------------ backward parts
// do_e = dM_e / a_ii;  dE = do x wo;  per block partial sums of do_e E_e and do_e
__global__ void __launch_bounds__( 256 ) dec_bwd_k( int64_t E, const double* __restrict__ dM,
                                                    const int32_t* __restrict__ erow, const double* __restrict__ diag,
                                                    const uint16_t* __restrict__ X, const float* __restrict__ wo,
                                                    float* __restrict__ dE, float* __restrict__ part )
{
   __shared__ float sdo[ 256 ];
   const int64_t e0 = (int64_t) blockIdx.x * 256;
   const int64_t e = e0 + threadIdx.x;
   float d = 0.f;
   if( e < E )
      d = (float) ( dM[ e ] / diag[ erow[ e ] ] );
   sdo[ threadIdx.x ] = d;
   __syncthreads();
   const int64_t cnt = min( (int64_t) 256, E - e0 );
   for( int64_t q = threadIdx.x; q < cnt * kH; q += 256 ) {
      const int64_t r = e0 + q / kH;
      const int c = (int) ( q % kH );
      dE[ r * kH + c ] = sdo[ q / kH ] * wo[ c ];
   }
   if( threadIdx.x < kH + 1 ) {
      float s = 0.f;
      for( int64_t r = 0; r < cnt; r++ )
         s += threadIdx.x < kH ? sdo[ r ] * fb( X[ ( e0 + r ) * kH + threadIdx.x ] ) : sdo[ r ];
      part[ (int64_t) blockIdx.x * ( kH + 1 ) + threadIdx.x ] = s;
   }
}

// dT = BF16( dH * [H > 0] )
__global__ void mask_k( int64_t cnt, const float* __restrict__ dX, const uint16_t* __restrict__ Y,
                        uint16_t* __restrict__ out )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q < cnt )
      out[ q ] = bf( fb( Y[ q ] ) > 0.f ? dX[ q ] : 0.f );
}

// dS = BF16( (dE + dG[row] / deg[row]) * [E > 0] )
__global__ void edge_bwd_k( int64_t E, const float* __restrict__ dE, const float* __restrict__ dG,
                            const int32_t* __restrict__ erow, const int32_t* __restrict__ deg,
                            const uint16_t* __restrict__ Y, uint16_t* __restrict__ dS )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= E * kH )
      return;
   const int64_t e = q / kH;
   const int c = (int) ( q % kH );
   const int32_t i = erow[ e ];
   const float g = dE[ q ] + dG[ (int64_t) i * kH + c ] / (float) deg[ i ];
   dS[ q ] = bf( fb( Y[ q ] ) > 0.f ? g : 0.f );
}

// dQ_i = BF16( sum over row i of dS ),  dP_j = BF16( sum over column j of dS )
__global__ void scatter_k( int64_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ cptr,
                           const int32_t* __restrict__ cperm, const uint16_t* __restrict__ dS,
                           uint16_t* __restrict__ dP, uint16_t* __restrict__ dQ )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= n * kH )
      return;
   const int64_t i = q / kH;
   const int c = (int) ( q % kH );
   float sq = 0.f, sp = 0.f;
   for( int32_t e = off[ i ]; e < off[ i + 1 ]; e++ )
      sq += fb( dS[ (int64_t) e * kH + c ] );
   for( int32_t p = cptr[ i ]; p < cptr[ i + 1 ]; p++ )
      sp += fb( dS[ (int64_t) cperm[ p ] * kH + c ] );
   dQ[ q ] = bf( sq );
   dP[ q ] = bf( sp );
}

// encoder gradients: per block partials of x^T (dX * [Y > 0]) (4 x 64) and
// of the column sums (64)
__global__ void __launch_bounds__( 256 ) enc_bwd_k( int64_t R, const float* __restrict__ x,
                                                    const float* __restrict__ dX, const uint16_t* __restrict__ Y,
                                                    float* __restrict__ part )
{
   const int64_t r0 = (int64_t) blockIdx.x * 256;
   const int64_t cnt = min( (int64_t) 256, R - r0 );
   const int c = threadIdx.x & 63, k = threadIdx.x >> 6;  // k = 0..3: input feature
   float s = 0.f, sb = 0.f;
   for( int64_t r = r0; r < r0 + cnt; r++ ) {
      const float g = fb( Y[ r * kH + c ] ) > 0.f ? dX[ r * kH + c ] : 0.f;
      s += x[ r * 4 + k ] * g;
      sb += g;
   }
   float* out = part + (int64_t) blockIdx.x * ( 5 * kH );
   out[ k * kH + c ] = s;
   if( k == 0 )
      out[ 4 * kH + c ] = sb;
}

// ------------------------------------------------------------ parameters
__global__ void refresh_bf_k( const float* __restrict__ params, uint16_t* __restrict__ W, uint16_t* __restrict__ WT )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= (int64_t) kL * kMats * kH * kH )
      return;
   const int64_t mat = q / ( kH * kH );
   const int l = (int) ( mat / kMats ), m = (int) ( mat % kMats );
   const int k = (int) ( ( q / kH ) % kH ), nn = (int) ( q % kH );
   const uint16_t v = bf( params[ oW( l, m ) + k * kH + nn ] );
   W[ mat * kH * kH + k * kH + nn ] = v;   // [k][n]: Bn for X W^T
   WT[ mat * kH * kH + nn * kH + k ] = v;  // [n][k]: Bn for X W
}

__global__ void adam_k( int64_t cnt, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                        float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2 )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= cnt )
      return;
   const float gq = g[ q ];
   const float mq = b1 * m[ q ] + ( 1.f - b1 ) * gq;
   const float vq = b2 * v[ q ] + ( 1.f - b2 ) * gq * gq;
   m[ q ] = mq;
   v[ q ] = vq;
   p[ q ] -= lr * ( mq / c1 ) / ( sqrtf( vq / c2 ) + eps );
}

inline dim3 g1( int64_t cnt, int b = 256 )
{
   return dim3( (unsigned) std::max< int64_t >( 1, ceil_div( cnt, b ) ) );
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

struct tnlmi_spai
{
   DeviceArray< float > params, grads, m, v;
   DeviceArray< uint16_t > W, WT;  // BF16 copies of the 15 64 x 64 matrices
   int64_t step = 0;
};

namespace tnlmi {
namespace {

const uint16_t* Wn( const tnlmi_spai& S, int l, int m )  // Bn for X W^T
{
   return S.W.ptr + (int64_t) ( l * kMats + m ) * kH * kH;
}
const uint16_t* WTn( const tnlmi_spai& S, int l, int m )  // Bn for X W
{
   return S.WT.ptr + (int64_t) ( l * kMats + m ) * kH * kH;
}

void refresh( tnlmi_spai& S, hipStream_t s )
{
   hipLaunchKernelGGL( refresh_bf_k, g1( (int64_t) kL * kMats * kH * kH ), dim3( 256 ), 0, s, S.params.ptr, S.W.ptr,
                       S.WT.ptr );
}

void check_csr( const tnlmi_matrix* A )
{
   if( ! A || ! A->segments )
      fail( TNLMI_INVALID_ARGUMENT, "spai: NULL matrix" );
   if( A->segments->format != TNLMI_CSR || A->value_type != TNLMI_F64 )
      fail( TNLMI_INVALID_ARGUMENT, "spai: the matrix must be CSR with FP64 values" );
   if( A->matrix_type == TNLMI_SYMMETRIC )
      fail( TNLMI_NOT_IMPLEMENTED, "spai: a general matrix is needed (the network reads both triangles' entries)" );
   if( A->rows != A->columns || A->rows <= 0 )
      fail( TNLMI_INVALID_ARGUMENT, "spai: the matrix must be square and not empty" );
}

void build_graph( Graph& g, const tnlmi_matrix& A, hipStream_t s )
{
   g.n = A.rows;
   g.E = A.segments->elements;
   const int32_t* off = A.segments->offsets.ptr;
   const int32_t* cols = A.column_indexes;
   const double* vals = (const double*) A.values;
   g.erow.allocate( g.E );
   g.deg.allocate( g.n );
   g.diag.allocate( g.n );
   g.isd.allocate( g.E );
   g.xe.allocate( g.E * 4 );
   g.xn.allocate( g.n * 4 );
   hipLaunchKernelGGL( g_rows, g1( g.n ), dim3( 256 ), 0, s, g.n, off, cols, vals, g.erow.ptr, g.deg.ptr,
                       g.diag.ptr, g.isd.ptr );
   hipLaunchKernelGGL( g_features, g1( g.n ), dim3( 256 ), 0, s, g.n, off, cols, vals, g.diag.ptr, g.xe.ptr,
                       g.xn.ptr );
   // edges grouped by column (stable radix sort by column): the transposed scatter
   DeviceArray< int32_t > ids, skeys;
   ids.allocate( g.E );
   skeys.allocate( g.E );
   g.cperm.allocate( g.E );
   g.cptr.allocate( g.n + 1 );
   hipLaunchKernelGGL( iota_k, g1( g.E ), dim3( 256 ), 0, s, g.E, ids.ptr );
   size_t bytes = 0;
   TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( nullptr, bytes, cols, skeys.ptr, ids.ptr, g.cperm.ptr, (int) g.E,
                                                  0, 32, s ) );
   DeviceArray< char > tmp;
   tmp.allocate( (int64_t) bytes + 1 );
   TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( tmp.ptr, bytes, cols, skeys.ptr, ids.ptr, g.cperm.ptr, (int) g.E,
                                                  0, 32, s ) );
   hipLaunchKernelGGL( g_colptr, g1( g.E + 1 ), dim3( 256 ), 0, s, g.E, g.n, skeys.ptr, g.cptr.ptr );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );  // the temporaries go out of scope
}

struct Acts
{
   DeviceArray< uint16_t > E[ kL + 1 ], H[ kL + 1 ], G[ kL + 1 ];
   DeviceArray< float > P, Q;
};

template< int NT, int EPI >
void gemm( int64_t R, Term t0, Term t1, Term t2, const EpiArgs& e, hipStream_t s )
{
   hipLaunchKernelGGL( ( gemm64< NT, EPI > ), g1( R, 64 ), dim3( 256 ), 0, s, R, t0, t1, t2, e );
}

// forward pass; keep = every layer's activations (training), else only the last
void forward( const tnlmi_spai& S, const Graph& g, const int32_t* off, const int32_t* cols, Acts& a, bool keep,
              double* M, hipStream_t s )
{
   const int64_t n = g.n, E = g.E;
   const float* p = S.params.ptr;
   const int nE = keep ? kL + 1 : 2, nH = keep ? kL + 1 : 2;
   for( int i = 0; i < nE; i++ )
      a.E[ i ].allocate( E * kH );
   for( int i = 0; i < nH; i++ ) {
      a.H[ i ].allocate( n * kH );
      a.G[ i ].allocate( n * kH );
   }
   a.P.allocate( n * kH );
   a.Q.allocate( n * kH );
   hipLaunchKernelGGL( encode_k, g1( E * kH ), dim3( 256 ), 0, s, E, g.xe.ptr, p + oAe, p + oae, a.E[ 0 ].ptr );
   hipLaunchKernelGGL( encode_k, g1( n * kH ), dim3( 256 ), 0, s, n, g.xn.ptr, p + oAn, p + oan, a.H[ 0 ].ptr );
   const Term none{ nullptr, nullptr };
   for( int l = 0; l < kL; l++ ) {
      const int ci = keep ? l : ( l & 1 ), ni = keep ? l + 1 : ( ( l + 1 ) & 1 );
      EpiArgs ep{};
      ep.yf = a.P.ptr;
      gemm< 1, EPI_STORE >( n, Term{ a.H[ ci ].ptr, WTn( S, l, mWu ) }, none, none, ep, s );
      ep.yf = a.Q.ptr;
      gemm< 1, EPI_STORE >( n, Term{ a.H[ ci ].ptr, WTn( S, l, mWv ) }, none, none, ep, s );
      EpiArgs ee{};
      ee.yb = a.E[ ni ].ptr;
      ee.bias = p + obe( l );
      ee.P = a.P.ptr;
      ee.Q = a.Q.ptr;
      ee.col = cols;
      ee.row = g.erow.ptr;
      gemm< 1, EPI_EDGE >( E, Term{ a.E[ ci ].ptr, WTn( S, l, mWe ) }, none, none, ee, s );
      hipLaunchKernelGGL( mean_k, g1( n * kH ), dim3( 256 ), 0, s, n, off, a.E[ ni ].ptr, a.G[ ni ].ptr );
      EpiArgs en{};
      en.yb = a.H[ ni ].ptr;
      en.bias = p + obh( l );
      gemm< 2, EPI_NODE >( n, Term{ a.H[ ci ].ptr, WTn( S, l, mWh ) }, Term{ a.G[ ni ].ptr, WTn( S, l, mWg ) }, none,
                           en, s );
   }
   const int last = keep ? kL : ( kL & 1 );
   hipLaunchKernelGGL( decode_k, g1( E ), dim3( 256 ), 0, s, E, a.E[ last ].ptr, p + owo, p + obo, g.erow.ptr,
                       cols, g.isd.ptr, g.diag.ptr, M );
   TNLMI_HIP( hipGetLastError() );
}

// out[e] = sum of part[b * width + e] over the nblk blocks: chunks of
// consecutive blocks (reduce_parts_chunks), then the chunk sums in order;
// work holds kRedChunks * width floats
void reduce_partials( const float* part, int64_t nblk, int width, float* out, float* work, hipStream_t s )
{
   const int64_t L = ceil_div( std::max< int64_t >( nblk, 1 ), kRedChunks );
   const int64_t chunks = std::max< int64_t >( 1, ceil_div( nblk, L ) );
   hipLaunchKernelGGL( reduce_parts_chunks, dim3( (unsigned) ceil_div( width, 64 ), (unsigned) chunks ), dim3( 64 ),
                       0, s, part, nblk, width, L, work );
   hipLaunchKernelGGL( reduce_parts, g1( width ), dim3( 256 ), 0, s, work, chunks, width, out );
}

void wgrad( int64_t R, const uint16_t* X, const uint16_t* dY, float* gW, float* gb, DeviceArray< float >& part,
            DeviceArray< float >& bpart, float* work, hipStream_t s )
{
   const int64_t nb = ceil_div( R, kWRows );
   hipLaunchKernelGGL( wgrad64, dim3( (unsigned) nb ), dim3( 256 ), 0, s, R, X, dY, part.ptr,
                       gb ? bpart.ptr : (float*) nullptr );
   reduce_partials( part.ptr, nb, kH * kH, gW, work, s );
   if( gb )
      reduce_partials( bpart.ptr, nb, kH, gb, work, s );
}

// probe loss on the device: returns L (host) and dL/dM per edge in dM
double probe_loss( const tnlmi_matrix& A, const Graph& g, const double* M, const double* Z, int K, double* dM,
                   hipStream_t s )
{
   const int64_t n = g.n, E = g.E;
   const int32_t* off = A.segments->offsets.ptr;
   const int32_t* cols = A.column_indexes;
   DeviceArray< double > Y, Rk, part, out;
   Y.allocate( n * K );
   Rk.allocate( n * K );
   const int64_t nb = ceil_div( n, 256 );
   part.allocate( nb * K );
   out.allocate( 1 );
   for( int k = 0; k < K; k++ ) {
      hipLaunchKernelGGL( csr_mv, g1( n ), dim3( 256 ), 0, s, n, off, cols, (const double*) A.values, Z + k * n,
                          Y.ptr + k * n );
      hipLaunchKernelGGL( csr_mv, g1( n ), dim3( 256 ), 0, s, n, off, cols, M, Y.ptr + k * n, Rk.ptr + k * n );
      hipLaunchKernelGGL( resid_k, dim3( (unsigned) nb ), dim3( 256 ), 0, s, n, Z + k * n, Rk.ptr + k * n,
                          part.ptr + k * nb );
   }
   hipLaunchKernelGGL( sum_k, dim3( 1 ), dim3( 256 ), 0, s, part.ptr, nb * K, 1.0 / ( (double) K * (double) n ),
                       out.ptr );
   if( dM )
      hipLaunchKernelGGL( dM_k, g1( E ), dim3( 256 ), 0, s, E, n, K, g.erow.ptr, cols, Rk.ptr, Y.ptr, dM );
   double L = 0;
   TNLMI_HIP( hipMemcpyAsync( &L, out.ptr, sizeof( double ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return L;
}

// gradients of the loss into S.grads
double loss_grad( tnlmi_spai& S, const tnlmi_matrix& A, const double* Z, int K, double* Mout, hipStream_t s )
{
   Graph g;
   build_graph( g, A, s );
   const int64_t n = g.n, E = g.E;
   const int32_t* off = A.segments->offsets.ptr;
   const int32_t* cols = A.column_indexes;
   Acts a;
   DeviceArray< double > M, dM;
   M.allocate( E );
   dM.allocate( E );
   forward( S, g, off, cols, a, true, M.ptr, s );
   const double L = probe_loss( A, g, M.ptr, Z, K, dM.ptr, s );
   if( Mout )
      TNLMI_HIP( hipMemcpyAsync( Mout, M.ptr, sizeof( double ) * E, hipMemcpyDeviceToDevice, s ) );
   float* gr = S.grads.ptr;
   const float* p = S.params.ptr;
   DeviceArray< float > dE, dEn, dH, dG, part, bpart, dpart, rwork;
   DeviceArray< uint16_t > dS, dT, dP, dQ;
   dE.allocate( E * kH );
   dEn.allocate( E * kH );
   dH.allocate( n * kH );
   dG.allocate( n * kH );
   dS.allocate( E * kH );
   dT.allocate( n * kH );
   dP.allocate( n * kH );
   dQ.allocate( n * kH );
   const int64_t nbE = ceil_div( E, kWRows );
   part.allocate( nbE * kH * kH );
   bpart.allocate( nbE * kH );
   dpart.allocate( ceil_div( E, 256 ) * ( 5 * kH ) + 1 );
   rwork.allocate( (int64_t) kRedChunks * kH * kH );  // chunk sums of reduce_partials (widest: kH * kH)
   // decoder
   const int64_t nb256 = ceil_div( E, 256 );
   hipLaunchKernelGGL( dec_bwd_k, dim3( (unsigned) nb256 ), dim3( 256 ), 0, s, E, dM.ptr, g.erow.ptr, g.diag.ptr,
                       a.E[ kL ].ptr, p + owo, dE.ptr, dpart.ptr );
   reduce_partials( dpart.ptr, nb256, kH + 1, gr + owo, rwork.ptr, s );
   TNLMI_HIP( hipMemsetAsync( dH.ptr, 0, sizeof( float ) * n * kH, s ) );
   const Term none{ nullptr, nullptr };
   for( int l = kL - 1; l >= 0; l-- ) {
      // node update: dT = dH * [H > 0]
      hipLaunchKernelGGL( mask_k, g1( n * kH ), dim3( 256 ), 0, s, n * kH, dH.ptr, a.H[ l + 1 ].ptr, dT.ptr );
      wgrad( n, a.H[ l ].ptr, dT.ptr, gr + oW( l, mWh ), gr + obh( l ), part, bpart, rwork.ptr, s );
      wgrad( n, a.G[ l + 1 ].ptr, dT.ptr, gr + oW( l, mWg ), nullptr, part, bpart, rwork.ptr, s );
      EpiArgs eg{};
      eg.yf = dG.ptr;
      gemm< 1, EPI_STORE >( n, Term{ dT.ptr, Wn( S, l, mWg ) }, none, none, eg, s );
      // edge update: dS = (dE + dG[row] / deg) * [E > 0]
      hipLaunchKernelGGL( edge_bwd_k, g1( E * kH ), dim3( 256 ), 0, s, E, dE.ptr, dG.ptr, g.erow.ptr, g.deg.ptr,
                          a.E[ l + 1 ].ptr, dS.ptr );
      wgrad( E, a.E[ l ].ptr, dS.ptr, gr + oW( l, mWe ), gr + obe( l ), part, bpart, rwork.ptr, s );
      EpiArgs ed{};
      ed.yf = dEn.ptr;
      gemm< 1, EPI_STORE >( E, Term{ dS.ptr, Wn( S, l, mWe ) }, none, none, ed, s );
      hipLaunchKernelGGL( scatter_k, g1( n * kH ), dim3( 256 ), 0, s, n, off, g.cptr.ptr, g.cperm.ptr, dS.ptr, dP.ptr,
                          dQ.ptr );
      wgrad( n, a.H[ l ].ptr, dP.ptr, gr + oW( l, mWu ), nullptr, part, bpart, rwork.ptr, s );
      wgrad( n, a.H[ l ].ptr, dQ.ptr, gr + oW( l, mWv ), nullptr, part, bpart, rwork.ptr, s );
      EpiArgs eh{};
      eh.yf = dH.ptr;
      gemm< 3, EPI_STORE >( n, Term{ dT.ptr, Wn( S, l, mWh ) }, Term{ dP.ptr, Wn( S, l, mWu ) },
                            Term{ dQ.ptr, Wn( S, l, mWv ) }, eh, s );
      std::swap( dE.ptr, dEn.ptr );
   }
   // encoder
   hipLaunchKernelGGL( enc_bwd_k, dim3( (unsigned) nb256 ), dim3( 256 ), 0, s, E, g.xe.ptr, dE.ptr, a.E[ 0 ].ptr,
                       dpart.ptr );
   reduce_partials( dpart.ptr, nb256, 5 * kH, gr + oAe, rwork.ptr, s );
   const int64_t nbn = ceil_div( n, 256 );
   hipLaunchKernelGGL( enc_bwd_k, dim3( (unsigned) nbn ), dim3( 256 ), 0, s, n, g.xn.ptr, dH.ptr, a.H[ 0 ].ptr,
                       dpart.ptr );
   reduce_partials( dpart.ptr, nbn, 5 * kH, gr + oAn, rwork.ptr, s );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return L;
}

}  // namespace
}  // namespace tnlmi

extern "C" {

int64_t tnlmi_spai_parameter_count( void )
{
   return kParams;
}

tnlmi_status tnlmi_spai_create( const float* h_params, tnlmi_spai** out )
{
   return guarded( [ & ] {
      if( ! out || ! h_params )
         fail( TNLMI_INVALID_ARGUMENT, "spai_create: NULL argument" );
      *out = nullptr;
      auto* S = new tnlmi_spai;
      try {
         for( auto* a : { &S->params, &S->grads, &S->m, &S->v } )
            a->allocate( kParams );
         S->W.allocate( (int64_t) kL * kMats * kH * kH );
         S->WT.allocate( (int64_t) kL * kMats * kH * kH );
         TNLMI_HIP( hipMemcpy( S->params.ptr, h_params, sizeof( float ) * kParams, hipMemcpyHostToDevice ) );
         TNLMI_HIP( hipMemset( S->grads.ptr, 0, sizeof( float ) * kParams ) );
         TNLMI_HIP( hipMemset( S->m.ptr, 0, sizeof( float ) * kParams ) );
         TNLMI_HIP( hipMemset( S->v.ptr, 0, sizeof( float ) * kParams ) );
         refresh( *S, nullptr );
         TNLMI_HIP( hipDeviceSynchronize() );
      }
      catch( ... ) {
         delete S;
         throw;
      }
      *out = S;
   } );
}

void tnlmi_spai_destroy( tnlmi_spai* S )
{
   delete S;
}

tnlmi_status tnlmi_spai_get_parameters( const tnlmi_spai* S, float* h_params, float* h_grads )
{
   return guarded( [ & ] {
      if( ! S )
         fail( TNLMI_INVALID_ARGUMENT, "spai: NULL handle" );
      if( h_params )
         TNLMI_HIP( hipMemcpy( h_params, S->params.ptr, sizeof( float ) * kParams, hipMemcpyDeviceToHost ) );
      if( h_grads )
         TNLMI_HIP( hipMemcpy( h_grads, S->grads.ptr, sizeof( float ) * kParams, hipMemcpyDeviceToHost ) );
   } );
}

tnlmi_status tnlmi_spai_apply( const tnlmi_spai* S, const tnlmi_matrix* A, double* d_M, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! S || ! d_M )
         fail( TNLMI_INVALID_ARGUMENT, "spai_apply: NULL argument" );
      check_csr( A );
      hipStream_t s = as_stream( stream );
      Graph g;
      build_graph( g, *A, s );
      Acts a;
      forward( *S, g, A->segments->offsets.ptr, A->column_indexes, a, false, d_M, s );
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

tnlmi_status tnlmi_spai_loss_grad( tnlmi_spai* S, const tnlmi_matrix* A, const double* d_Z, int32_t K, double* loss,
                                   double* d_M, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! S || ! d_Z || K <= 0 || ! loss )
         fail( TNLMI_INVALID_ARGUMENT, "spai_loss_grad: invalid argument" );
      check_csr( A );
      *loss = loss_grad( *S, *A, d_Z, K, d_M, as_stream( stream ) );
   } );
}

tnlmi_status tnlmi_spai_train_step( tnlmi_spai* S, const tnlmi_matrix* A, int32_t K, uint64_t probe_seed,
                                    float learning_rate, double* loss, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! S || K <= 0 || ! loss )
         fail( TNLMI_INVALID_ARGUMENT, "spai_train_step: invalid argument" );
      check_csr( A );
      hipStream_t s = as_stream( stream );
      DeviceArray< double > Z;
      Z.allocate( (int64_t) K * A->rows );
      hipLaunchKernelGGL( rademacher_k, g1( (int64_t) K * A->rows ), dim3( 256 ), 0, s, (int64_t) K * A->rows,
                          probe_seed, Z.ptr );
      *loss = loss_grad( *S, *A, Z.ptr, K, nullptr, s );
      S->step++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow( b1, (float) S->step ), c2 = 1.f - std::pow( b2, (float) S->step );
      hipLaunchKernelGGL( adam_k, g1( kParams ), dim3( 256 ), 0, s, kParams, S->params.ptr, S->grads.ptr, S->m.ptr,
                          S->v.ptr, learning_rate, b1, b2, 1e-8f, c1, c2 );
      refresh( *S, s );
      TNLMI_HIP( hipGetLastError() );
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

tnlmi_status tnlmi_bf16_gemm64( int64_t rows, const uint16_t* d_X, const uint16_t* d_W, float* d_Y,
                                tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( rows < 0 || ! d_X || ! d_W || ! d_Y )
         fail( TNLMI_INVALID_ARGUMENT, "bf16_gemm64: invalid argument" );
      hipStream_t s = as_stream( stream );
      // Y = X W with W row-major [k][n]: the kernel reads B as Bn[n][k], the transpose
      DeviceArray< uint16_t > WT;
      WT.allocate( kH * kH );
      std::vector< uint16_t > h( kH * kH ), t( kH * kH );
      TNLMI_HIP( hipMemcpyAsync( h.data(), d_W, sizeof( uint16_t ) * kH * kH, hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      for( int k = 0; k < kH; k++ )
         for( int nn = 0; nn < kH; nn++ )
            t[ nn * kH + k ] = h[ k * kH + nn ];
      TNLMI_HIP( hipMemcpyAsync( WT.ptr, t.data(), sizeof( uint16_t ) * kH * kH, hipMemcpyHostToDevice, s ) );
      EpiArgs e{};
      e.yf = d_Y;
      if( rows > 0 )
         gemm< 1, EPI_STORE >( rows, Term{ d_X, WT.ptr }, Term{ nullptr, nullptr }, Term{ nullptr, nullptr }, e, s );
      TNLMI_HIP( hipGetLastError() );
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

}  // extern "C"
