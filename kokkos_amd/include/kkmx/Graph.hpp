#ifndef KKMX_GRAPH_HPP
#define KKMX_GRAPH_HPP
#include <kkmx/Parallel.hpp>
#endif
