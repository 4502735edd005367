// Internal helpers shared by the host and device translation units.
#ifndef OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_
#define OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_

#include <string>

namespace octlb {

// Records msg as the calling thread's last error and returns code.
int SetError(int code, const std::string& msg);

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_OCTLB_INTERNAL_H_
