// SayakaTreeData.h -- the reference header name of this part of the call surface (a driver written for
// the reference includes it unchanged); the declarations are in tree_data.h.
#pragma once
#include "tree_data.h"
