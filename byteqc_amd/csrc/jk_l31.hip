// Instantiates the J/K kernels whose bra shell pair has angular momenta (3, 1).
#include "jk_launch.cuh"
namespace mjk {
template hipError_t launch_lab<3, 1>(int, int, bool, const JKArgs &, hipStream_t, long *);
}
