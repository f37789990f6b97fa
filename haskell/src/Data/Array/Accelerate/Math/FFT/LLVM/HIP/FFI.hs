{-# LANGUAGE ForeignFunctionInterface #-}
-- |
-- Module      : Data.Array.Accelerate.Math.FFT.LLVM.HIP.FFI
-- License     : BSD3
-- Stability   : experimental
--
-- Bindings of libafft's C ABI (include/afft.h), the MI355X FFT engine behind the HIP plugin.
-- This is the part of the plugin that replaces the cuFFT binding (Foreign.CUDA.FFT, used at
-- LLVM/PTX.hs:112-132): every entry point takes raw device pointers, a raw hipStream_t and plain
-- integers, and returns an int status (0 = success, text from afft_strerror).
--
-- Type mapping (C -> Haskell): int -> CInt, int64_t -> Int64, double -> CDouble,
-- const int64_t * -> Ptr Int64, void * and const void * (device pointers, the stream) -> Ptr (),
-- const char * -> CString.
--
-- The front-end calls are 'safe': the first call for a shape builds its plan (device allocations
-- and table generation, milliseconds), which must not block the other Haskell threads of the
-- capability.  Later calls only enqueue kernels.
--
module Data.Array.Accelerate.Math.FFT.LLVM.HIP.FFI (

  -- * Status
  afft_strerror,

  -- * The front-end operations (FFT.hs:63-173)
  afft_fft,
  afft_fft1D,
  afft_fft2D,
  afft_fft3D,
  afft_fft2D_many,

  -- * Centring fused around a transform (DFT/Centre.hs)
  afft_fft_centred,

  -- * Adjoints and the spectral convolution
  afft_fft_vjp,
  afft_spectral_conv_fwd,
  afft_spectral_conv_vjp,
  afft_spectral_conv_vjp_residual,

  -- * Codes
  modeForward, modeReverse, modeInverse,
  dtypeC64, dtypeC128,
  opFFT, opFFT1D, opFFT2D, opFFT3D,

) where

import Data.Int                                                     ( Int64 )
import Foreign.C.String                                             ( CString )
import Foreign.C.Types                                              ( CDouble(..), CInt(..) )
import Foreign.Ptr                                                  ( Ptr )


-- Mode (AFFT_FORWARD / AFFT_REVERSE / AFFT_INVERSE), element type (AFFT_C64 / AFFT_C128), and
-- the operation codes of afft_fft_vjp.
modeForward, modeReverse, modeInverse :: CInt
modeForward = 0
modeReverse = 1
modeInverse = 2

dtypeC64, dtypeC128 :: CInt
dtypeC64  = 0
dtypeC128 = 1

opFFT, opFFT1D, opFFT2D, opFFT3D :: CInt
opFFT   = 0
opFFT1D = 1
opFFT2D = 2
opFFT3D = 3


foreign import ccall unsafe "afft_strerror"
  afft_strerror :: CInt -> IO CString

-- int afft_fft(int mode, int dtype, int rank, const int64_t *extents, const void *d_in,
--              void *d_out, void *stream)
foreign import ccall safe "afft_fft"
  afft_fft :: CInt -> CInt -> CInt -> Ptr Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft1D(int mode, int dtype, int64_t n, const void *d_in, void *d_out, void *stream)
foreign import ccall safe "afft_fft1D"
  afft_fft1D :: CInt -> CInt -> Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft2D(int mode, int dtype, int64_t h, int64_t w, const void *d_in, void *d_out,
--                void *stream)
foreign import ccall safe "afft_fft2D"
  afft_fft2D :: CInt -> CInt -> Int64 -> Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft3D(int mode, int dtype, int64_t d, int64_t h, int64_t w, const void *d_in,
--                void *d_out, void *stream)
foreign import ccall safe "afft_fft3D"
  afft_fft3D :: CInt -> CInt -> Int64 -> Int64 -> Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft2D_many(int mode, int dtype, int64_t batch, int64_t h, int64_t w,
--                     const void *d_in, void *d_out, void *stream)
foreign import ccall safe "afft_fft2D_many"
  afft_fft2D_many :: CInt -> CInt -> Int64 -> Int64 -> Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft_centred(int mode, int dtype, int rank, const int64_t *extents, int pre, int post,
--                      const void *d_in, void *d_out, void *stream)
foreign import ccall safe "afft_fft_centred"
  afft_fft_centred :: CInt -> CInt -> CInt -> Ptr Int64 -> CInt -> CInt -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_fft_vjp(int op, int mode, int dtype, int rank, const int64_t *extents,
--                  const void *d_gy, void *d_gx, void *stream)
foreign import ccall safe "afft_fft_vjp"
  afft_fft_vjp :: CInt -> CInt -> CInt -> CInt -> Ptr Int64 -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_spectral_conv_fwd(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
--                            const void *d_H, void *d_y, void *stream)
foreign import ccall safe "afft_spectral_conv_fwd"
  afft_spectral_conv_fwd :: CInt -> Int64 -> Int64 -> Int64 -> Ptr () -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_spectral_conv_vjp(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
--                            const void *d_H, const void *d_gy, void *d_gx, void *d_gH,
--                            void *stream)
foreign import ccall safe "afft_spectral_conv_vjp"
  afft_spectral_conv_vjp :: CInt -> Int64 -> Int64 -> Int64 -> Ptr () -> Ptr () -> Ptr () -> Ptr () -> Ptr () -> Ptr () -> IO CInt

-- int afft_spectral_conv_vjp_residual(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
--                                     int x_is_spectrum, const void *d_H, const void *d_y,
--                                     const void *d_t, double alpha, void *d_gx, void *d_gH,
--                                     void *stream)
foreign import ccall safe "afft_spectral_conv_vjp_residual"
  afft_spectral_conv_vjp_residual :: CInt -> Int64 -> Int64 -> Int64 -> Ptr () -> CInt -> Ptr () -> Ptr () -> Ptr () -> CDouble -> Ptr () -> Ptr () -> Ptr () -> IO CInt
