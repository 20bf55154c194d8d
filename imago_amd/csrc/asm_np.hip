// asm_np.hip - k_assemble instances for the NP integral kind.
#include "assemble_launch.hpp"

namespace imago {
hipError_t launch_assemble_np(const AssembleLaunch &L) { return launch_kind<K_NP>(L); }
}  // namespace imago
