// FP64 instantiation.  Built with -ffp-contract=off: with no fused
// multiply-adds the results equal the reference CPU path bit for bit.
#include "launch_impl.hpp"

template struct xfd::Launch<double>;
template struct xfd::BatchLaunch<double>;
template struct xfd::AdjLaunch<double>;
template struct xfd::StepLaunch<double>;
