// asm_ke.hip - k_assemble instances for the KE integral kind.
#include "assemble_launch.hpp"

namespace imago {
hipError_t launch_assemble_ke(const AssembleLaunch &L) { return launch_kind<K_KE>(L); }
}  // namespace imago
