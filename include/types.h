/* types.h -- integer types, Fourier-prime record and status codes of the
 * CUMODP C API.  Same names, widths and values as the reference
 * (include/types.h:45-53, :110-118, :123-146) so callers compile unchanged.
 * The reference selects the word types with a platform macro
 * (LINUXINTEL64 etc.); this library targets 64-bit Linux only, where the
 * reference's choice is sfixn = int. */
#ifndef CUMODP_AMD_TYPES_H
#define CUMODP_AMD_TYPES_H

/* symbols of the shared library are hidden unless marked */
#ifndef CUMODP_API
#define CUMODP_API __attribute__((visibility("default")))
#endif

#ifndef __modpn_integer_types__
#define __modpn_integer_types__
typedef int sfixn;
typedef long longfixnum;
typedef unsigned int usfixn;
typedef unsigned long ulongfixnum;
typedef int int32;
typedef unsigned int uint32;
#endif

/* p = c * 2^npow + 1 */
typedef struct {
    sfixn p;
    sfixn c;
    sfixn npow;
    sfixn rpow;  /* 2^rpow is the smallest power of two larger than p */
    sfixn r;     /* 2^rpow mod p */
    sfixn rnpow; /* rpow - npow */
} fprime_t;

typedef enum {
    CUMODP_SUCCESS = 0x000,
    CUMODP_FAILURE = 0x001,
    CUMODP_ASSUMPTION_ERROR = 0x002,
    CUMODP_NO_CUDA_DEVICE = 0x100,
    CUMODP_HAS_CUDA_DEVICE = 0x101,
    CUMODP_FFT_ERROR = 0x200,
    CUMODP_FOURIER_DEGREE_TOO_SMALL = 0x201,
    CUMODP_FFT_SIZE_TOO_LARGE = 0x202,
    CUMODP_FFT_SIZE_TOO_SMALL = 0x203,
    CUMODP_SUBRES_ERROR = 0x300,
    CUMODP_SUBRES_COARSE_ERROR = 0x301,
    CUMODP_SUBRES_FINE_ERROR = 0x302,
    CUMODP_SUBRES_NON_REGULAR = 0x303,
    CUMODP_RES2_ERROR = 0x304,
    CUMODP_RES3_ERROR = 0x305,
    CUMODP_OUT_OF_MEMORY = 0x401,
    CUMODP_UNKNOWN_ERROR = 0x999
} cumodp_err;

#endif
