// asm_ep.hip - k_assemble instances for the EP integral kind.
#include "assemble_launch.hpp"

namespace imago {
hipError_t launch_assemble_ep(const AssembleLaunch &L) { return launch_ep<>(L); }
}  // namespace imago
