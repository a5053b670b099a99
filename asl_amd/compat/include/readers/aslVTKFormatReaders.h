// aslVTKFormatReaders.h — asl::readSurface (src/readers/aslVTKFormatReaders.h:
// 66-78) without VTK: the surface is read from an STL file (binary or ASCII)
// or a VTK XML PolyData file (.vtp: aslhip_vtp_read) and its level-set map is
// sampled on the GPU (aslhip_surface_distance_map).  The map is a float
// container whose ghost-inclusive grid is the sampled grid (makeData,
// src/utilities/aslVTKCasters.cxx:562-575): map <= 0 within 1.2 dx (Block
// variant) or 1.8 dx (dx variants) of the surface.
#ifndef ASLVTKFORMATREADERS_H
#define ASLVTKFORMATREADERS_H

#include <string>

#include "../aslDataInc.h"

namespace asl {

/// Grid: the surface's bounding box (grown by 1.25 % of its largest side, as
/// VTK's modeller does) sampled every dx.
SPDataWithGhostNodesACLData readSurface(const std::string &fileName, double dx,
                                        acl::CommandQueue queue = acl::hardware.defaultQueue);
/// Grid: the nodes of b (its ghost-inclusive grid is b; the internal block is b shrunk by one).
SPDataWithGhostNodesACLData readSurface(const std::string &fileName, Block &b,
                                        acl::CommandQueue queue = acl::hardware.defaultQueue);
/// As the dx variant, the box extended on each side by the given fractions of its extent.
SPDataWithGhostNodesACLData readSurface(const std::string &fileName, double dx, double offset_X0,
                                        double offset_XE, double offset_Y0, double offset_YE, double offset_Z0,
                                        double offset_ZE, acl::CommandQueue queue = acl::hardware.defaultQueue);

}  // namespace asl

#endif
