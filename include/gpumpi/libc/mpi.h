// `#include "mpi.h"` of a converted program.
#pragma once
#include "../mpi.h"
