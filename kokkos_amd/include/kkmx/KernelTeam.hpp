#ifndef KKMX_KERNELTEAM_HPP
#define KKMX_KERNELTEAM_HPP
#include <kkmx/Parallel.hpp>
#endif
