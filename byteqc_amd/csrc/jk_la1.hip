// Instantiates the J/K kernels whose bra shell has angular momentum 1.
#include "jk_launch.cuh"
namespace mjk {
template hipError_t launch_la<1>(int, int, int, bool, const JKArgs &, hipStream_t, long *);
}
