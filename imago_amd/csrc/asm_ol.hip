// asm_ol.hip - k_assemble instances for the OL integral kind.
#include "assemble_launch.hpp"

namespace imago {
hipError_t launch_assemble_ol(const AssembleLaunch &L) { return launch_kind<K_OL>(L); }
}  // namespace imago
