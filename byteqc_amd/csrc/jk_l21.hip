// Instantiates the J/K kernels whose bra shell pair has angular momenta (2, 1).
#include "jk_launch.cuh"
namespace mjk {
template hipError_t launch_lab<2, 1>(int, int, bool, const JKArgs &, hipStream_t, long *);
}
