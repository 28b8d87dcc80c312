/*
 * fft_api.h -- batched FFT plan / exec interface of the MI355X (gfx950) FFT engine.
 *
 * Drop-in for the reference header include/fft_api.h (namespace AsdSip, extern "C"):
 * same function names, enum values, argument order and status codes, so a C++
 * caller of the reference recompiles unchanged and links against libasdfft.so.
 * In C++ the out-parameters are references, exactly as in the reference; a C
 * caller (or an FFI such as ctypes) sees the same symbols with pointers, which
 * is the same ABI.
 *
 * Each declaration cites the reference declaration it replaces.
 *
 * Semantics (no normalisation in either direction; reference
 * docs/en/API_Reference/FFT/FFT_1D.md and the unit-test goldens):
 *   C2C FORWARD  X_k = sum x_n e^{-2 pi i nk/N}        C2C INVERSE  sum x_n e^{+2 pi i nk/N}
 *   R2C FORWARD  rfft (N reals -> N/2+1 bins)           R2C INVERSE  conj(rfft)
 *   C2R FORWARD  hfft(n=N)                              C2R INVERSE  unnormalised irfft(n=N)
 *   2D/3D        the same per axis; the real axis is the last one.
 * Tensors: 1D (batch, N) [HORIZONTAL] or (N, batch) [VERTICAL]; 2D (batch, X, Y);
 * 3D (batch, X, Y, Z); COMPLEX64 is interleaved (re, im) float32; data pointers
 * are HIP device pointers.  Exec is asynchronous on the bound HIP stream.
 * ASCEND_FFT_C2C_SEP: the C2C transform on a pair of FLOAT planes (real, imaginary),
 * 1D (HORIZONTAL, or VERTICAL of batch 1) and 3D plans; 2D plans reject it as the
 * reference does.
 */
#ifndef ASDSIP_FFT_API_H
#define ASDSIP_FFT_API_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#include "aspb_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aclTensor aclTensor;

#ifdef __cplusplus
namespace AsdSip {
#define ASDFFT_OUT(T) T&
#define ASDFFT_DEFAULT(v) = v
#else
#define ASDFFT_OUT(T) T*
#define ASDFFT_DEFAULT(v)
#endif

/* include/fft_api.h:22 */
typedef void* asdFftHandle;

/* include/fft_api.h:24 */
enum asdFftStatus { SUCCESS = 0, FAILED = 1 };

/* include/fft_api.h:26-34 */
enum asdFftType {
    ASCEND_FFT_C2C = 0x10,
    ASCEND_FFT_C2R = 0x11,
    ASCEND_FFT_R2C = 0x12,
    ASCEND_STFT_C2C = 0x20,
    ASCEND_STFT_C2R = 0x21,
    ASCEND_STFT_R2C = 0x22,
    ASCEND_FFT_C2C_SEP = 0x30
};

/* include/fft_api.h:36-39 */
enum asdFftDirection {
    ASCEND_FFT_FORWARD = 0x10,
    ASCEND_FFT_INVERSE = 0x11,
};

/* include/fft_api.h:41-44 */
enum asdFft1dDimType {
    ASCEND_FFT_HORIZONTAL = 0x10,
    ASCEND_FFT_VERTICAL = 0x11,
};

#ifndef __cplusplus
typedef enum asdFftType asdFftType;
typedef enum asdFftDirection asdFftDirection;
typedef enum asdFft1dDimType asdFft1dDimType;
#endif

/* Creates an empty plan handle.  Replaces include/fft_api.h:48 (impl core/fft/fft_api.cpp:1047-1065). */
AspbStatus asdFftCreate(ASDFFT_OUT(asdFftHandle) handle);

/* Binds a hipStream_t; NULL is rejected.  Replaces include/fft_api.h:51 (fft_api.cpp:1067-1082). */
AspbStatus asdFftSetStream(asdFftHandle handle, void* stream);

/* 1D plan; a handle can be initialised once.  Replaces include/fft_api.h:55-57 (fft_api.cpp:1124-1190). */
AspbStatus asdFftMakePlan1D(asdFftHandle handle, int64_t fftSize, asdFftType fftType, asdFftDirection direction,
                            int64_t batchSize, asdFft1dDimType dimType ASDFFT_DEFAULT(ASCEND_FFT_HORIZONTAL));

/* ISTFT plan for a complex64 (channel, nFft, nFrames) spectrum: center = true, normalized = false,
 * onesided = false, lengthOpt = 0, returnComplex = true, the reference's defaults (other values are
 * replaced with a warning, as the reference does).  Requires nFrames > 1, 0 < hopLengthOpt <= winLengthOpt == nFft,
 * nFft < 1500, hopLengthOpt < 1500 (else ACL_ERROR_INVALID_PARAM); an initialised handle ->
 * ACL_ERROR_INTERNAL_ERROR.  The workspace holds the inverse-transformed frames.
 * Replaces include/fft_api.h:59-62 (core/fft/istft.cpp:343-433, 525-570). */
AspbStatus asdFftIstftMakePlan(asdFftHandle handle, const aclTensor* input, const int64_t nFft,
                               const int64_t hopLengthOpt, const int64_t winLengthOpt,
                               const bool center ASDFFT_DEFAULT(true), const bool normalized ASDFFT_DEFAULT(false),
                               const bool onesidedOpt ASDFFT_DEFAULT(false), int64_t lengthOpt ASDFFT_DEFAULT(0),
                               const bool returnComplex ASDFFT_DEFAULT(true));

/* Replaces include/fft_api.h:65 (fft_api.cpp:1752-1788). */
AspbStatus asdFftExecC2C(asdFftHandle handle, const aclTensor* input, const aclTensor* output);

/* C2C of a plan made with ASCEND_FFT_C2C_SEP (1D, or 3D) on separated real / imaginary
 * FLOAT planes of one shape, (batch, N) or (batch, X, Y, Z).  Checks in the reference's
 * order: shapes differ -> ACL_ERROR_INTERNAL_ERROR; a non-FLOAT plane ->
 * ACL_ERROR_OP_INPUT_NOT_MATCH; plan not initialised or shape not the plan's ->
 * ACL_ERROR_INVALID_PARAM; null data, or a plan of another type -> ACL_ERROR_INTERNAL_ERROR.
 * outputReal/outputImag may equal inputReal/inputImag (in place).
 * Replaces include/fft_api.h:68-69 (fft_api.cpp:1846-1985). */
AspbStatus asdFftExecC2CSeparated(asdFftHandle handle, const aclTensor* inputReal, const aclTensor* inputImag,
                                  const aclTensor* outputReal, const aclTensor* outputImag);

/* Replaces include/fft_api.h:72 (fft_api.cpp:1790-1816). */
AspbStatus asdFftExecC2R(asdFftHandle handle, const aclTensor* input, const aclTensor* output);

/* Replaces include/fft_api.h:75 (fft_api.cpp:1818-1844). */
AspbStatus asdFftExecR2C(asdFftHandle handle, const aclTensor* input, const aclTensor* output);

/* out[b][i] = sum_t w[n_t] ifft(spec[b][:, t])[n_t] / sum_t w[n_t]^2 with n_t = nFft/2 + i - t*hop
 * (overlap-add of the windowed inverse frames, squared-window envelope, centre crop); output
 * complex64 (channel, nFft + hop*(nFrames-1) - 2*(nFft/2)), window float (nFft).  Shape or dtype
 * not the plan's -> ACL_ERROR_INVALID_PARAM; null data -> ACL_ERROR_INTERNAL_ERROR; the output
 * must not overlap the input or window (no in-place update).
 * Replaces include/fft_api.h:77-78 (core/fft/istft.cpp:224-316, 437-523, 572-595). */
AspbStatus asdFftExecIstft(asdFftHandle handle, const aclTensor* input, const aclTensor* windowOpt,
                           const aclTensor* output);

/* Frees the plan and its device tables.  Replaces include/fft_api.h:81 (fft_api.cpp:1234-1243). */
AspbStatus asdFftDestroy(asdFftHandle handle);

/* Replaces include/fft_api.h:83-84 (fft_api.cpp:1245-1289). */
AspbStatus asdFftMakePlan2D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, asdFftType fftType,
                            asdFftDirection direction, int32_t batchSize);

/* Replaces include/fft_api.h:86-87 (fft_api.cpp:1291-1337). */
AspbStatus asdFftMakePlan3D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, int64_t fftSizeZ,
                            asdFftType fftType, asdFftDirection direction, int32_t batchSize);

/* Bytes of caller-owned device workspace the plan needs (0 if none); also records the
 * amount for the exec-time guard.  Replaces include/fft_api.h:89 (fft_api.cpp:1339-1360). */
AspbStatus asdFftGetWorkspaceSize(asdFftHandle handle, ASDFFT_OUT(size_t) workspaceSize);

/* Replaces include/fft_api.h:91 (fft_api.cpp:1362-1374). */
AspbStatus asdFftSetWorkspace(asdFftHandle handle, void* workspace);

/* hipStreamSynchronize on the bound stream.  Replaces include/fft_api.h:93 (fft_api.cpp:1376-1390). */
AspbStatus asdFftSynchronize(asdFftHandle handle);

/* Replaces include/fft_api.h:95 (fft_api.cpp:1392-1404). */
AspbStatus asdFftGetType(asdFftHandle handle, ASDFFT_OUT(asdFftType) fftType);

/* Batched 2D circular shift of complex64 slices on `stream` (a hipStream_t):
 *   asdFftShift : out[s][(r + ceil(row/2)) % row][(c + ceil(col/2)) % col] = in[s*dist + r*col + c]
 *   asdIfftShift: out[s][(r + row/2) % row][(c + col/2) % col] = float(scale) * in[s*dist + r*col + c]
 * dist <= 0 means contiguous slices (dist = row * col); the output is a contiguous
 * (batch, row, col) tensor.  workspace is unused (the reference takes none either).
 * Replaces include/fft_api.h:105-106 and :121-122 (core/fft/fftshift.cpp:101-138, ifftshift.cpp). */
AspbStatus asdFftShift(int64_t row, int64_t col, int64_t batch, int64_t dist, const aclTensor* input,
                       aclTensor* output, void* stream, void* workspace ASDFFT_DEFAULT(nullptr));
AspbStatus asdIfftShift(int64_t row, int64_t col, int64_t batch, int64_t dist, double scale, const aclTensor* input,
                        aclTensor* output, void* stream, void* workspace ASDFFT_DEFAULT(nullptr));

#ifdef __cplusplus
}  // namespace AsdSip
#endif

#undef ASDFFT_OUT
#undef ASDFFT_DEFAULT

#ifdef __cplusplus
}
#endif

#endif
