// Device unistd.h subset for converted programs: nothing beyond the standard header.
#pragma once
#include <unistd.h>
