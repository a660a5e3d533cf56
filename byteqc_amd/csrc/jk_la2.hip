// Instantiates the J/K kernels whose bra shell has angular momentum 2.
#include "jk_launch.cuh"
namespace mjk {
template hipError_t launch_la<2>(int, int, int, bool, const JKArgs &, hipStream_t, long *);
}
