// The reference's PrintUtil.h provides LogLevel (PrintUtil.h:33-43); the shim
// keeps the enum in Algo.h.
#ifndef SDL_AMD_PRINTUTIL_H
#define SDL_AMD_PRINTUTIL_H
#include "Algo.h"
#endif
