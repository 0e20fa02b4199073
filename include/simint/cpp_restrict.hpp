/* C++ spelling of C99 `restrict` (skel/simint/cpp_restrict.hpp:7). */
#pragma once
#define restrict __restrict__
