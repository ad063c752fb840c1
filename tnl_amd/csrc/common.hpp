// common.hpp -- shared helpers of the gfx950 engine: status/error plumbing,
// numeric type traits, wave64 reductions and XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "tnlmi.h"

namespace tnlmi {

// ---------------------------------------------------------------- errors
// Errors travel as C++ exceptions inside the library and are converted to a
// status + thread-local message at every extern "C" boundary (guarded()).
struct Error : std::runtime_error
{
   tnlmi_status status;
   Error( tnlmi_status s, const std::string& m ) : std::runtime_error( m ), status( s ) {}
};

void set_last_error( const std::string& m );

template< typename F >
tnlmi_status guarded( F&& f ) noexcept
{
   try {
      f();
      return TNLMI_OK;
   }
   catch( const Error& e ) {
      set_last_error( e.what() );
      return e.status;
   }
   catch( const std::bad_alloc& e ) {
      set_last_error( e.what() );
      return TNLMI_BAD_ALLOC;
   }
   catch( const std::exception& e ) {
      set_last_error( e.what() );
      return TNLMI_RUNTIME_ERROR;
   }
   catch( ... ) {
      set_last_error( "unknown error" );
      return TNLMI_RUNTIME_ERROR;
   }
}

[[noreturn]] inline void fail( tnlmi_status s, const std::string& m )
{
   throw Error( s, m );
}

inline void hip_check( hipError_t e, const char* what )
{
   if( e != hipSuccess )
      fail( e == hipErrorOutOfMemory ? TNLMI_BAD_ALLOC : TNLMI_BACKEND_ERROR,
            std::string( what ) + ": " + hipGetErrorString( e ) );
}
#define TNLMI_HIP( call ) ::tnlmi::hip_check( ( call ), #call )

// Device buffer owned by the library (segment arrays, solver work vectors).
template< typename T >
struct DeviceArray
{
   T* ptr = nullptr;
   int64_t size = 0;
   DeviceArray() = default;
   DeviceArray( const DeviceArray& ) = delete;
   DeviceArray& operator=( const DeviceArray& ) = delete;
   DeviceArray( DeviceArray&& o ) noexcept : ptr( o.ptr ), size( o.size )
   {
      o.ptr = nullptr;
      o.size = 0;
   }
   DeviceArray& operator=( DeviceArray&& o ) noexcept
   {
      std::swap( ptr, o.ptr );
      std::swap( size, o.size );
      return *this;
   }
   ~DeviceArray() { reset(); }
   void reset()
   {
      if( ptr )
         (void) hipFree( ptr );
      ptr = nullptr;
      size = 0;
   }
   void allocate( int64_t n )
   {
      if( n == size && ptr )
         return;
      reset();
      if( n > 0 )
         TNLMI_HIP( hipMalloc( (void**) &ptr, sizeof( T ) * (size_t) n ) );
      size = n;
   }
};

inline hipStream_t as_stream( tnlmi_stream s )
{
   return reinterpret_cast< hipStream_t >( s );
}

// ------------------------------------------------------------ numerics
struct bf16
{
   uint16_t bits;
};

__host__ __device__ inline float bf16_to_float( uint16_t b )
{
   uint32_t u = (uint32_t) b << 16;
   float f;
   memcpy( &f, &u, 4 );
   return f;
}
// round to nearest, ties to even; NaN stays NaN
__host__ __device__ inline uint16_t float_to_bf16( float f )
{
   uint32_t u;
   memcpy( &u, &f, 4 );
   if( ( u & 0x7f800000u ) == 0x7f800000u && ( u & 0x007fffffu ) )
      return (uint16_t) ( ( u >> 16 ) | 0x40 );
   const uint32_t lsb = ( u >> 16 ) & 1u;
   u += 0x7fffu + lsb;
   return (uint16_t) ( u >> 16 );
}

template< typename T >
struct Load
{
   __device__ static T get( const T* p, int64_t i ) { return p[ i ]; }
};

// Value-type load with conversion to the compute type.
template< typename V, typename C >
__device__ __forceinline__ C load_val( const V* p, int64_t i )
{
   return (C) p[ i ];
}
template<>
__device__ __forceinline__ float load_val< bf16, float >( const bf16* p, int64_t i )
{
   return bf16_to_float( p[ i ].bits );
}
template<>
__device__ __forceinline__ double load_val< bf16, double >( const bf16* p, int64_t i )
{
   return (double) bf16_to_float( p[ i ].bits );
}

// Streaming (non-temporal) loads for matrix data that is read exactly once,
// so that the x vector keeps the L2 / MALL capacity.
template< typename T >
__device__ __forceinline__ T nt_load( const T* p )
{
   return __builtin_nontemporal_load( p );
}
template< typename V, typename C >
__device__ __forceinline__ C nt_load_val( const V* p, int64_t i )
{
   return (C) nt_load( p + i );
}
template<>
__device__ __forceinline__ float nt_load_val< bf16, float >( const bf16* p, int64_t i )
{
   return bf16_to_float( nt_load( reinterpret_cast< const uint16_t* >( p ) + i ) );
}
template<>
__device__ __forceinline__ double nt_load_val< bf16, double >( const bf16* p, int64_t i )
{
   return (double) bf16_to_float( nt_load( reinterpret_cast< const uint16_t* >( p ) + i ) );
}

// ------------------------------------------------------ wave64 reductions
constexpr int kWave = 64;

// Sum over groups of G consecutive lanes (G power of two <= 64); every lane of
// the group receives the group sum.
template< int G, typename T >
__device__ __forceinline__ T group_sum( T v )
{
#pragma unroll
   for( int o = G / 2; o > 0; o /= 2 )
      v += __shfl_xor( v, o, G );
   return v;
}
template< typename T >
__device__ __forceinline__ T group_sum_dyn( T v, int g )
{
   for( int o = g / 2; o > 0; o /= 2 )
      v += __shfl_xor( v, o, 64 );
   return v;
}

// The value of the wavefront's first active lane, as a value the compiler
// knows to be uniform: loops and branches on it stay scalar (no per-lane exit
// masks around barriers).  For values that are equal in every lane.
__device__ __forceinline__ double wave_uniform( double v )
{
   const uint64_t b = (uint64_t) __double_as_longlong( v );
   const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane( (int) (uint32_t) b );
   const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane( (int) (uint32_t) ( b >> 32 ) );
   return __longlong_as_double( (long long) ( ( (uint64_t) hi << 32 ) | lo ) );
}
__device__ __forceinline__ float wave_uniform( float v )
{
   return __int_as_float( __builtin_amdgcn_readfirstlane( __float_as_int( v ) ) );
}
__device__ __forceinline__ int32_t wave_uniform( int32_t v )
{
   return __builtin_amdgcn_readfirstlane( v );
}
__device__ __forceinline__ int64_t wave_uniform( int64_t v )
{
   const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane( (int) (uint32_t) v );
   const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane( (int) (uint32_t) ( (uint64_t) v >> 32 ) );
   return (int64_t) ( ( (uint64_t) hi << 32 ) | lo );
}

// Block-wide sum for blockDim.x == B (multiple of 64); result valid (and
// uniform) in all threads; all threads must call it.  smem must hold B/64 values.
template< int B, typename T >
__device__ __forceinline__ T block_sum( T v, T* smem )
{
   v = group_sum< 64 >( v );
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   __syncthreads();
   if( lane == 0 )
      smem[ w ] = v;
   __syncthreads();
   T r = 0;
#pragma unroll
   for( int i = 0; i < B / 64; i++ )
      r += smem[ i ];
   return wave_uniform( r );
}

// ------------------------------------------------------- XCD-aware mapping
// The dispatcher places workgroup b on XCD (b mod 8).  Renumbering so that each
// XCD walks one contiguous range of tiles keeps neighbouring rows (and the x
// entries they share) in one XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_contiguous( int64_t b, int64_t grid )
{
   const int64_t q = grid / kXcds, rem = grid % kXcds;
   const int64_t x = b % kXcds, j = b / kXcds;
   return x * q + ( x < rem ? x : rem ) + j;
}

inline int64_t ceil_div( int64_t a, int64_t b )
{
   return ( a + b - 1 ) / b;
}

}  // namespace tnlmi
