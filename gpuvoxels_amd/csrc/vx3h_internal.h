// vx3h_internal.h -- declarations shared by the engine's translation units.
#ifndef VX3H_INTERNAL_H
#define VX3H_INTERNAL_H
#include <string>
namespace vx3 {
extern thread_local std::string g_err;   // last error of the calling thread (vx3h_last_error)
}
#endif
