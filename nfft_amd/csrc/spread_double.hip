// B / B^T kernels instantiated for double (kernel templates in spread_impl.hpp).
#include "spread_impl.hpp"

namespace nfft_hip {
template void launch_interp<double>(const cplx<double> *, cplx<double> *, const Engine<double> &, hipStream_t);
template void launch_spread<double>(const cplx<double> *, cplx<double> *, const Engine<double> &, hipStream_t);
template void launch_gather_tiles<double>(const Engine<double> &, cplx<double> *, hipStream_t);
}  // namespace nfft_hip
