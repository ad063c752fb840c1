// runtime.cpp -- error state and version of the C ABI.
#include <string>

#include "common.hpp"

namespace tnlmi {

namespace {
thread_local std::string last_error;
}

void set_last_error( const std::string& m )
{
   last_error = m;
}

}  // namespace tnlmi

extern "C" {

const char* tnlmi_last_error( void )
{
   return tnlmi::last_error.c_str();
}

const char* tnlmi_version( void )
{
   return "tnlmi 0.1 (gfx950, wave64)";
}

}
