// SayakaBoundaryPatch.h -- the reference header name of this part of the call surface (a driver written for
// the reference includes it unchanged); the declarations are in fill_patch.h.
#pragma once
#include "fill_patch.h"
