// Passed through for converted programs: the standard header works in HIP device code
// as it is (HIP provides the device math overloads and assert).
#pragma once
#include <stddef.h>
