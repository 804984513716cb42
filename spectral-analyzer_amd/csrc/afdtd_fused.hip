// afdtd_fused.hip — placeholder until the fused sweep lands.
#include "afdtd_common.cuh"
bool fused_supported(const AcousticFDTDState* st) { (void)st; return false; }
hipError_t launch_fused_step(AcousticFDTDState* st) { (void)st; return hipErrorInvalidValue; }
