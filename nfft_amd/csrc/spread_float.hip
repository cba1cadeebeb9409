// B / B^T kernels instantiated for float (kernel templates in spread_impl.hpp).
#include "spread_impl.hpp"

namespace nfft_hip {
template void launch_interp<float>(const cplx<float> *, cplx<float> *, const Engine<float> &, hipStream_t, int,
                                 const BinGeom *);
template void launch_spread<float>(const cplx<float> *, cplx<float> *, const Engine<float> &, hipStream_t, int,
                                 const BinGeom *, cplx<float> *);
template void launch_gather_tiles<float>(const Engine<float> &, cplx<float> *, hipStream_t, int,
                                       const BinGeom *, cplx<float> *);
template bool batch_kernels<float>(const Engine<float> &);
}  // namespace nfft_hip
