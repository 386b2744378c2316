// Error translation for the C ABI: every entry point runs its body through
// hg_guard, which maps exceptions to status codes and records the message.
#pragma once

#include <exception>
#include <new>

#include "hyteg_amd.h"
#include "runtime.hpp"

template < class F >
inline int hg_guard( hg_ctx* h, F&& body )
{
   if ( !h )
      return HG_ERR_ARG;
   auto* ctx = reinterpret_cast< hg::Context* >( h );
   try
   {
      HG_CHECK_HIP( hipSetDevice( ctx->device ) );
      body( ctx );
      ctx->err.clear();
      return HG_OK;
   } catch ( const hg::Error& e )
   {
      ctx->err = e.what();
      return e.code;
   } catch ( const std::bad_alloc& )
   {
      ctx->err = "host out of memory";
      return HG_ERR_NOMEM;
   } catch ( const std::invalid_argument& e )
   {
      ctx->err = e.what();
      return HG_ERR_ARG;
   } catch ( const std::exception& e )
   {
      ctx->err = e.what();
      return HG_ERR_STATE;
   }
}
