// FP32 instantiation (fields, psi and coefficients in float; the TFSF and
// lumped corrections are evaluated in double and rounded once per record).
#include "launch_impl.hpp"

template struct xfd::Launch<float>;
template struct xfd::BatchLaunch<float>;
template struct xfd::AdjLaunch<float>;
template struct xfd::StepLaunch<float>;
