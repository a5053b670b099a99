// aclTypes.h — element type ids of ASL's acl layer (src/acl/aclTypes.h).
#ifndef ACLTYPES_H
#define ACLTYPES_H

namespace acl {

enum TypeID { TYPE_INT, TYPE_UINT, TYPE_FLOAT, TYPE_DOUBLE, TYPE_LONG };

template <typename T>
inline TypeID typeToTypeID();
template <>
inline TypeID typeToTypeID<int>() { return TYPE_INT; }
template <>
inline TypeID typeToTypeID<unsigned int>() { return TYPE_UINT; }
template <>
inline TypeID typeToTypeID<float>() { return TYPE_FLOAT; }
template <>
inline TypeID typeToTypeID<double>() { return TYPE_DOUBLE; }
template <>
inline TypeID typeToTypeID<long>() { return TYPE_LONG; }

// The OpenCL queue of ASL (src/acl/aclHardware.h:141-163).  Work of each
// lattice runs on its own HIP stream; the queue argument is accepted and
// ignored.
struct CommandQueue {};
struct Hardware {
  CommandQueue defaultQueue;
};
extern Hardware hardware;

}  // namespace acl

#endif
