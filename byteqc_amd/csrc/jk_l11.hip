// Instantiates the J/K kernels whose bra shell pair has angular momenta (1, 1).
#include "jk_launch.cuh"
namespace mjk {
template hipError_t launch_lab<1, 1>(int, int, bool, const JKArgs &, hipStream_t, long *);
}
