#include "cd.hpp"
#define double cd
#include "../o_rad.c"
