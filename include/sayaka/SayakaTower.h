// SayakaTower.h -- the reference header name of this part of the call surface (a driver written for
// the reference includes it unchanged); the declarations are in mac_solver.h.
#pragma once
#include "mac_solver.h"
