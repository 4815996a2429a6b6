// The POA kernels compiled again with wide aligned sets (kMaxAlignedWide
// entries per node instead of kMaxAligned) and "_wide" kernel and launcher
// names. The host reruns here the windows whose graph columns hold more
// distinct bytes than the default kernel's aligned sets allow (status
// kPoaAlignedFull), e.g. reads with IUPAC ambiguity codes.
#define RH_POA_WIDE 1
#define RH_POA_ALIGNED kMaxAlignedWide
#include "poa_kernel.hip"
