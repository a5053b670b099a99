// aslGeomInc.h — distance-function generators and CSG (src/aslGeomInc.h:54-150).
#ifndef ASLGEOMINC_H
#define ASLGEOMINC_H

#include <vector>

#include "acl/aclTypes.h"
#include "data/aslDataWithGhostNodes.h"
#include "math/aslDistanceFunction.h"

namespace asl {

SPDistanceFunction generateDFSphere(double r, const AVec<double> &c);
SPDistanceFunction generateDFCylinderInf(double r, const AVec<double> &l, const AVec<double> &c);
SPDistanceFunction generateDFCylinder(double r, const AVec<double> &l, const AVec<double> &c);
SPDistanceFunction generateDFCone(double r, const AVec<double> &l, const AVec<double> &a);
SPDistanceFunction generateDFPlane(const AVec<double> &n, const AVec<double> &p0);
SPDistanceFunction generateDFConvexPolygonPrism(std::vector<AVec<double>> points);
SPDistanceFunction generateDFConvexPolygonPyramid(std::vector<AVec<double>> points, AVec<double> a);
SPDistanceFunction generateDFInBlock(const Block &b, unsigned int nG);
SPDistanceFunction operator&(SPDistanceFunction a, SPDistanceFunction b);
SPDistanceFunction operator|(SPDistanceFunction a, SPDistanceFunction b);
SPDistanceFunction operator-(SPDistanceFunction a);
SPDistanceFunction normalize(SPDistanceFunction a, double dx);

// The map of df sampled on the ghost-inclusive block (src/aslGeomInc.cxx:187-221).
SPAbstractDataWithGhostNodes generateDataContainer_SP(const Block &b, SPDistanceFunction df, unsigned int gN,
                                                      acl::TypeID t);
void initData(SPAbstractDataWithGhostNodes d, SPDistanceFunction f);

}  // namespace asl

#endif
