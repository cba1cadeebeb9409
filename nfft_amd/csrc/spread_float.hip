// B / B^T kernels instantiated for float (kernel templates in spread_impl.hpp).
#include "spread_impl.hpp"

namespace nfft_hip {
template void launch_interp<float>(const cplx<float> *, cplx<float> *, const Engine<float> &, hipStream_t);
template void launch_spread<float>(const cplx<float> *, cplx<float> *, const Engine<float> &, hipStream_t);
template void launch_gather_tiles<float>(const Engine<float> &, cplx<float> *, hipStream_t);
}  // namespace nfft_hip
