#include "cd.hpp"
#define double cd
#include "../o_cell.c"
