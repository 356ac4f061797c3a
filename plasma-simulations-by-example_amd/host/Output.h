// Output: VTK ImageData fields, screen line and runtime_diags.csv, with the reference's
// file names, array names and column order (ch3/ver2/Output.h, Output.cpp) so that its
// ParaView states and plotting scripts read these files unchanged.
#ifndef PICAMD_HOST_OUTPUT_H
#define PICAMD_HOST_OUTPUT_H

#include <vector>

#include "Species.h"
#include "World.h"

namespace Output {
void fields(World &world, std::vector<Species> &species);
void screenOutput(World &world, std::vector<Species> &species);
void diagOutput(World &world, std::vector<Species> &species);
}  // namespace Output

#endif
