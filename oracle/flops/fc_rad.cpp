#include "cd.hpp"
#define O_FLOPCOUNT 1
#define double cd
#include "../o_rad.c"
