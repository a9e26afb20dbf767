// Device sys/time.h for converted programs: gettimeofday from the GPU wall clock.
#pragma once
#include "../time.h"
