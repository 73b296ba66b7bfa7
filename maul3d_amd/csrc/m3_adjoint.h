// placeholder
#ifndef M3_ADJOINT_H
#define M3_ADJOINT_H
#include "m3_serial.h"
#endif
