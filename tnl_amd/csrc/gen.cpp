// gen.cpp -- seeded host-side generators of the benchmark matrices and
// vectors (counter-based random numbers, OpenMP over rows).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace {

inline uint64_t splitmix( uint64_t z )
{
   z += 0x9e3779b97f4a7c15ull;
   z = ( z ^ ( z >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
   z = ( z ^ ( z >> 27 ) ) * 0x94d049bb133111ebull;
   return z ^ ( z >> 31 );
}
inline uint64_t rnd( uint64_t seed, uint64_t stream, uint64_t i )
{
   return splitmix( splitmix( seed ^ splitmix( stream + 0x632be59bd9b4e019ull ) ) + i );
}
inline double unit( uint64_t u )
{
   return (double) ( u >> 11 ) * ( 1.0 / 9007199254740992.0 );
}

}  // namespace

extern "C" {

int64_t tnlmi_gen_laplace2d( int64_t nx, int64_t ny, int32_t* off, int32_t* col, double* val )
{
   const int64_t n = nx * ny;
   const int64_t nnz = 5 * n - 2 * nx - 2 * ny;
   if( ! off )
      return nnz;
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const int64_t ix = i % nx, iy = i / nx;
      // entries of the grid lines above: 5*nx - 2 each, nx fewer on line 0
      int64_t p = iy * ( 5 * nx - 2 ) - ( iy > 0 ? nx : 0 );
      // entries of this line left of ix
      const int64_t per = 5 - ( iy == 0 ) - ( iy == ny - 1 );
      p += ix * per - ( ix > 0 ? 1 : 0 );
      off[ i ] = (int32_t) p;
      int64_t q = p;
      if( iy > 0 ) {
         col[ q ] = (int32_t) ( i - nx );
         val[ q++ ] = -1.0;
      }
      if( ix > 0 ) {
         col[ q ] = (int32_t) ( i - 1 );
         val[ q++ ] = -1.0;
      }
      col[ q ] = (int32_t) i;
      val[ q++ ] = 4.0;
      if( ix < nx - 1 ) {
         col[ q ] = (int32_t) ( i + 1 );
         val[ q++ ] = -1.0;
      }
      if( iy < ny - 1 ) {
         col[ q ] = (int32_t) ( i + nx );
         val[ q++ ] = -1.0;
      }
   }
   off[ n ] = (int32_t) nnz;
   return nnz;
}

int64_t tnlmi_gen_poisson3d( int64_t nx, int64_t ny, int64_t nz, int32_t* off, int32_t* col, double* val )
{
   const int64_t n = nx * ny * nz, pl = nx * ny;
   const int64_t nnz = 7 * n - 2 * ( ny * nz + nx * nz + nx * ny );
   if( ! off )
      return nnz;
   // per-row counts, then a serial scan, then the fill
   std::vector< int32_t > cnt( n );
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const int64_t ix = i % nx, iy = ( i / nx ) % ny, iz = i / pl;
      cnt[ i ] = 7 - ( ix == 0 ) - ( ix == nx - 1 ) - ( iy == 0 ) - ( iy == ny - 1 ) - ( iz == 0 ) - ( iz == nz - 1 );
   }
   int64_t acc = 0;
   for( int64_t i = 0; i < n; i++ ) {
      off[ i ] = (int32_t) acc;
      acc += cnt[ i ];
   }
   off[ n ] = (int32_t) acc;
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const int64_t ix = i % nx, iy = ( i / nx ) % ny, iz = i / pl;
      int64_t q = off[ i ];
      auto put = [ & ]( int64_t c, double v ) {
         col[ q ] = (int32_t) c;
         val[ q++ ] = v;
      };
      if( iz > 0 )
         put( i - pl, -1.0 );
      if( iy > 0 )
         put( i - nx, -1.0 );
      if( ix > 0 )
         put( i - 1, -1.0 );
      put( i, 6.0 );
      if( ix < nx - 1 )
         put( i + 1, -1.0 );
      if( iy < ny - 1 )
         put( i + nx, -1.0 );
      if( iz < nz - 1 )
         put( i + pl, -1.0 );
   }
   return nnz;
}

void tnlmi_gen_zipf_sizes( int64_t n, double s, int32_t kmax, double mean, int32_t lo, int32_t hi, uint64_t seed,
                           int32_t* sizes )
{
   std::vector< double > cdf( kmax );
   double z = 0;
   for( int32_t k = 1; k <= kmax; k++ )
      z += std::pow( (double) k, -s );
   double c = 0;
   for( int32_t k = 1; k <= kmax; k++ ) {
      c += std::pow( (double) k, -s ) / z;
      cdf[ k - 1 ] = c;
   }
   cdf[ kmax - 1 ] = 1.0;
   std::vector< int32_t > k( n );
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const double u = unit( rnd( seed, 1, i ) );
      k[ i ] = (int32_t) ( std::lower_bound( cdf.begin(), cdf.end(), u ) - cdf.begin() ) + 1;
   }
   const int64_t target = (int64_t) std::llround( mean * (double) n );
   auto total = [ & ]( double scale ) {
      int64_t t = 0;
#pragma omp parallel for reduction( + : t ) schedule( static )
      for( int64_t i = 0; i < n; i++ )
         t += std::min< int64_t >( hi, std::max< int64_t >( lo, (int64_t) std::floor( k[ i ] * scale + 0.5 ) ) );
      return t;
   };
   double a = 0.0, b = 2.0 * hi;
   for( int it = 0; it < 100; it++ ) {
      const double m = 0.5 * ( a + b );
      if( total( m ) <= target )
         a = m;
      else
         b = m;
   }
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ )
      sizes[ i ] = (int32_t) std::min< int64_t >( hi, std::max< int64_t >( lo, (int64_t) std::floor( k[ i ] * a + 0.5 ) ) );
   int64_t diff = target - total( a );
   for( int64_t pass = 0; diff != 0 && pass < 1000000; pass++ ) {
      bool moved = false;
      for( int64_t i = 0; i < n && diff != 0; i++ ) {
         if( diff > 0 && sizes[ i ] < hi ) {
            sizes[ i ]++;
            diff--;
            moved = true;
         }
         else if( diff < 0 && sizes[ i ] > lo ) {
            sizes[ i ]--;
            diff++;
            moved = true;
         }
      }
      if( ! moved )
         break;
   }
}

void tnlmi_gen_random_columns( int64_t n, int64_t ncols, const int32_t* off, uint64_t seed, int32_t* col )
{
#pragma omp parallel
   {
      std::vector< int32_t > buf;
#pragma omp for schedule( dynamic, 256 )
      for( int64_t r = 0; r < n; r++ ) {
         const int64_t L = std::min< int64_t >( off[ r + 1 ] - off[ r ], ncols );
         int32_t* out = col + off[ r ];
         if( L * 2 > ncols ) {
            // selection sampling (Knuth, TAOCP vol. 2, algorithm S)
            int64_t need = L, j = 0;
            for( int64_t c = 0; c < ncols && need > 0; c++ ) {
               const double u = unit( rnd( seed, 3 + 2 * (uint64_t) r, (uint64_t) c ) );
               if( (double) ( ncols - c ) * u < (double) need ) {
                  out[ j++ ] = (int32_t) c;
                  need--;
               }
            }
            continue;
         }
         buf.clear();
         uint64_t draw = 0;
         while( (int64_t) buf.size() < L ) {
            while( (int64_t) buf.size() < L )
               buf.push_back( (int32_t) ( rnd( seed, 2 + 2 * (uint64_t) r, draw++ ) % (uint64_t) ncols ) );
            std::sort( buf.begin(), buf.end() );
            buf.erase( std::unique( buf.begin(), buf.end() ), buf.end() );
         }
         std::copy( buf.begin(), buf.end(), out );
      }
   }
}

// Batch of 2D 5-point variable-coefficient diffusion systems (block-diagonal
// CSR).  System k: side s_k = round(exp(U[ln smin, ln smax])), cells
// (ix, iy), face coefficients c ~ LogNormal(0, 1) on every cell face
// (boundary faces included: Dirichlet), row = sum of the cell's 4 face
// coefficients on the diagonal, -c of each interior face off the diagonal.
// Pass NULL arrays to get the sizes only: sides[nsys] and *rows, *nnz.
void tnlmi_gen_batched_diffusion( int64_t nsys, int32_t smin, int32_t smax, uint64_t seed, int32_t* sides,
                                  int64_t* rows, int64_t* nnz, int32_t* offsets, int32_t* columns, double* values )
{
   std::vector< int64_t > row0( nsys + 1, 0 ), nz0( nsys + 1, 0 );
   for( int64_t k = 0; k < nsys; k++ ) {
      const double u = unit( rnd( seed, 101, k ) );
      const double ls = std::log( (double) smin ) + u * ( std::log( (double) smax ) - std::log( (double) smin ) );
      const int32_t s = (int32_t) std::min< double >( smax, std::max< double >( smin, std::floor( std::exp( ls ) + 0.5 ) ) );
      sides[ k ] = s;
      row0[ k + 1 ] = row0[ k ] + (int64_t) s * s;
      nz0[ k + 1 ] = nz0[ k ] + 5 * (int64_t) s * s - 4 * (int64_t) s;
   }
   *rows = row0[ nsys ];
   *nnz = nz0[ nsys ];
   if( ! offsets )
      return;
#pragma omp parallel for schedule( dynamic, 4 )
   for( int64_t k = 0; k < nsys; k++ ) {
      const int64_t s = sides[ k ];
      // horizontal faces h[iy][ix] between (ix-1, iy) and (ix, iy), ix = 0..s; vertical v[iy][ix], iy = 0..s
      auto hface = [ & ]( int64_t ix, int64_t iy ) {
         const double z = std::sqrt( -2.0 * std::log( 1.0 - unit( rnd( seed, 1000 + 2 * (uint64_t) k, 2 * ( iy * ( s + 1 ) + ix ) ) ) ) )
                          * std::cos( 6.283185307179586 * unit( rnd( seed, 1000 + 2 * (uint64_t) k, 2 * ( iy * ( s + 1 ) + ix ) + 1 ) ) );
         return std::exp( z );
      };
      auto vface = [ & ]( int64_t ix, int64_t iy ) {
         const double z = std::sqrt( -2.0 * std::log( 1.0 - unit( rnd( seed, 1001 + 2 * (uint64_t) k, 2 * ( iy * s + ix ) ) ) ) )
                          * std::cos( 6.283185307179586 * unit( rnd( seed, 1001 + 2 * (uint64_t) k, 2 * ( iy * s + ix ) + 1 ) ) );
         return std::exp( z );
      };
      int64_t q = nz0[ k ];
      for( int64_t iy = 0; iy < s; iy++ )
         for( int64_t ix = 0; ix < s; ix++ ) {
            const int64_t i = row0[ k ] + iy * s + ix;
            offsets[ i ] = (int32_t) q;
            const double w = hface( ix, iy ), e = hface( ix + 1, iy ), so = vface( ix, iy ), no = vface( ix, iy + 1 );
            if( iy > 0 ) {
               columns[ q ] = (int32_t) ( i - s );
               values[ q++ ] = -so;
            }
            if( ix > 0 ) {
               columns[ q ] = (int32_t) ( i - 1 );
               values[ q++ ] = -w;
            }
            columns[ q ] = (int32_t) i;
            values[ q++ ] = w + e + so + no;
            if( ix < s - 1 ) {
               columns[ q ] = (int32_t) ( i + 1 );
               values[ q++ ] = -e;
            }
            if( iy < s - 1 ) {
               columns[ q ] = (int32_t) ( i + s );
               values[ q++ ] = -no;
            }
         }
   }
   offsets[ row0[ nsys ] ] = (int32_t) nz0[ nsys ];
}

double tnlmi_gen_convdiff( int64_t m, uint64_t seed, int32_t* off, int32_t* col, double* val )
{
   // Peclet number and the eight stream-function modes of this instance
   const double pe = 1.0 + 99.0 * unit( rnd( seed, 201, 0 ) );
   double amp[ 8 ], kp[ 8 ], kq[ 8 ];
   for( int k = 0; k < 8; k++ ) {
      kp[ k ] = (double) ( 1 + rnd( seed, 202, 2 * k ) % 4 );
      kq[ k ] = (double) ( 1 + rnd( seed, 202, 2 * k + 1 ) % 4 );
      const double u1 = 1.0 - unit( rnd( seed, 203, 2 * k ) ), u2 = unit( rnd( seed, 203, 2 * k + 1 ) );
      amp[ k ] = std::sqrt( -2.0 * std::log( u1 ) ) * std::cos( 6.283185307179586 * u2 )
               / std::sqrt( kp[ k ] * kp[ k ] + kq[ k ] * kq[ k ] );
   }
   const double h = 1.0 / (double) ( m + 1 ), pi = 3.141592653589793;
   auto vel = [ & ]( int64_t ix, int64_t iy, double& vx, double& vy ) {
      const double x = h * (double) ( ix + 1 ), y = h * (double) ( iy + 1 );
      vx = 0.0;
      vy = 0.0;
      for( int k = 0; k < 8; k++ ) {
         const double sx = std::sin( pi * kp[ k ] * x ), cx = std::cos( pi * kp[ k ] * x );
         const double sy = std::sin( pi * kq[ k ] * y ), cy = std::cos( pi * kq[ k ] * y );
         vx += amp[ k ] * pi * kq[ k ] * sx * cy;   //  d psi / dy
         vy -= amp[ k ] * pi * kp[ k ] * cx * sy;   // -d psi / dx
      }
   };
   if( ! off )
      return pe;
   const int64_t n = m * m;
   double vmax = 0.0;
#pragma omp parallel for schedule( static ) reduction( max : vmax )
   for( int64_t i = 0; i < n; i++ ) {
      double vx, vy;
      vel( i % m, i / m, vx, vy );
      vmax = std::max( vmax, std::sqrt( vx * vx + vy * vy ) );
   }
   // rows scaled by h^2 / eps: diffusion stencil (4, -1), upwind convection h * Pe * v / max|v|
   const double c = vmax > 0.0 ? h * pe / vmax : 0.0;
   int64_t acc = 0;
   for( int64_t i = 0; i < n; i++ ) {
      const int64_t ix = i % m, iy = i / m;
      off[ i ] = (int32_t) acc;
      acc += 1 + ( ix > 0 ) + ( ix < m - 1 ) + ( iy > 0 ) + ( iy < m - 1 );
   }
   off[ n ] = (int32_t) acc;
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const int64_t ix = i % m, iy = i / m;
      double vx, vy;
      vel( ix, iy, vx, vy );
      vx *= c;
      vy *= c;
      int64_t q = off[ i ];
      if( iy > 0 ) {
         col[ q ] = (int32_t) ( i - m );
         val[ q++ ] = -1.0 - std::max( vy, 0.0 );
      }
      if( ix > 0 ) {
         col[ q ] = (int32_t) ( i - 1 );
         val[ q++ ] = -1.0 - std::max( vx, 0.0 );
      }
      col[ q ] = (int32_t) i;
      val[ q++ ] = 4.0 + std::fabs( vx ) + std::fabs( vy );
      if( ix < m - 1 ) {
         col[ q ] = (int32_t) ( i + 1 );
         val[ q++ ] = -1.0 - std::max( -vx, 0.0 );
      }
      if( iy < m - 1 ) {
         col[ q ] = (int32_t) ( i + m );
         val[ q++ ] = -1.0 - std::max( -vy, 0.0 );
      }
   }
   return pe;
}

void tnlmi_gen_uniform( int64_t n, double lo, double hi, uint64_t seed, uint64_t stream, double* out )
{
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ )
      out[ i ] = lo + ( hi - lo ) * unit( rnd( seed, stream, i ) );
}

void tnlmi_gen_normal( int64_t n, uint64_t seed, uint64_t stream, double* out )
{
   // Box-Muller on two independent uniforms per entry
#pragma omp parallel for schedule( static )
   for( int64_t i = 0; i < n; i++ ) {
      const double u1 = 1.0 - unit( rnd( seed, stream, 2 * (uint64_t) i ) );
      const double u2 = unit( rnd( seed, stream, 2 * (uint64_t) i + 1 ) );
      out[ i ] = std::sqrt( -2.0 * std::log( u1 ) ) * std::cos( 6.283185307179586 * u2 );
   }
}

}  // extern "C"
