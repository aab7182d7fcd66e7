// Shared C-ABI types (the public header lives in include/).
#pragma once
#include "spi_hip.h"

// Record the calling thread's last error message (reported by spi_last_error()).
void spi_set_error(const char* msg);
