#ifndef KKMX_KERNELMDRANGE_HPP
#define KKMX_KERNELMDRANGE_HPP
#include <kkmx/Parallel.hpp>
#endif
