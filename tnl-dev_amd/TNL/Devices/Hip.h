// Device tag of the MI355X path: the counterpart of TNL's Devices::Cuda (src/TNL/Devices/Cuda.h).
// Containers and matrices templated on it allocate with hipMalloc and run the tnlhip kernels.
#pragma once

#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "../../../include/tnlhip.h"

namespace TNL {
namespace Devices {

struct Hip
{
   static std::string getDeviceType() { return "Devices::Hip"; }
};

}  // namespace Devices

namespace Exceptions {

// Exceptions::CudaRuntimeError counterpart (src/TNL/Exceptions/CudaRuntimeError.h)
struct HipRuntimeError : std::runtime_error
{
   int status;
   HipRuntimeError( int s, const std::string& m ) : std::runtime_error( m ), status( s ) {}
};

}  // namespace Exceptions

namespace detail {

// Status codes of the C ABI to TNL's exception types.
inline void check( int status )
{
   if( status == TNLHIP_SUCCESS )
      return;
   const std::string msg = tnlhip_last_error();
   switch( status ) {
      case TNLHIP_ERROR_INVALID_ARGUMENT: throw std::invalid_argument( msg );
      case TNLHIP_ERROR_OUT_OF_MEMORY: throw std::bad_alloc();
      case TNLHIP_ERROR_CAPACITY: throw std::logic_error( msg );
      case TNLHIP_ERROR_BREAKDOWN: throw 1;  // GMRES.hpp:648-660 throws an int
      default: throw Exceptions::HipRuntimeError( status, msg );
   }
}

inline void checkHip( hipError_t e )
{
   if( e != hipSuccess )
      throw Exceptions::HipRuntimeError( TNLHIP_ERROR_RUNTIME, hipGetErrorString( e ) );
}

}  // namespace detail
}  // namespace TNL
