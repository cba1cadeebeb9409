// B / B^T kernels instantiated for double (kernel templates in spread_impl.hpp).
#include "spread_impl.hpp"

namespace nfft_hip {
template void launch_interp<double>(const cplx<double> *, cplx<double> *, const Engine<double> &, hipStream_t, int,
                                 const BinGeom *);
template void launch_spread<double>(const cplx<double> *, cplx<double> *, const Engine<double> &, hipStream_t, int,
                                 const BinGeom *, cplx<double> *);
template void launch_gather_tiles<double>(const Engine<double> &, cplx<double> *, hipStream_t, int,
                                       const BinGeom *, cplx<double> *);
template bool batch_kernels<double>(const Engine<double> &);
}  // namespace nfft_hip
