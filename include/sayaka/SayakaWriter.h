// SayakaWriter.h -- the reference header name of this part of the call surface (a driver written for
// the reference includes it unchanged); the declarations are in writer.h.
#pragma once
#include "writer.h"
