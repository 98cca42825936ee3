// SayakaBox.h -- the reference header name of this part of the call surface (a driver written for
// the reference includes it unchanged); the declarations are in index_box.h.
#pragma once
#include "index_box.h"
