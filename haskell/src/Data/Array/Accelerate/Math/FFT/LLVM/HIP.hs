{-# LANGUAGE GADTs               #-}
{-# LANGUAGE ScopedTypeVariables #-}
{-# LANGUAGE TypeApplications    #-}
{-# LANGUAGE TypeOperators       #-}
-- |
-- Module      : Data.Array.Accelerate.Math.FFT.LLVM.HIP
-- License     : BSD3
-- Stability   : experimental
--
-- FFT plugin for an AMD-GPU (HIP) backend of Accelerate, backed by libafft, the MI355X FFT
-- engine (include/afft.h).  It sits beside the LLVM.Native (FFTW) and LLVM.PTX (cuFFT) plugins
-- and exports the same four foreign functions with the same meaning (LLVM/PTX.hs:52-70):
--
--   * 'fft'    innermost axis of an array of rank 1..8, outer axes a batch (FFT.hs:63-84)
--   * 'fft1D', 'fft2D', 'fft3D'   every axis (FFT.hs:92-173)
--
-- As in the other plugins, Reverse and Inverse return the unnormalised transform and the
-- front-end divides (LLVM/PTX.hs:130-132, LLVM/Native/Base.hs:42-44).  The @Fused@ forms are
-- Inverse with the division done inside the engine's last pass: the front-end uses them with the
-- division moved into the fallback argument, so the other backends keep their semantics.
--
-- Extensions with no counterpart in the reference: a batched 'fft2DMany', the adjoint of each
-- transform ('fftVJP' ...), and the spectral convolution with its adjoint.
--
-- There is no plan cache here: libafft keeps its own, keyed on (device, element type, rank,
-- axes, extents) behind a mutex, and a plan holds no stream, so the setStream/exec race of the
-- cuFFT handle (LLVM/PTX.hs:121-124, PTX/Plans.hs:58-63) has no counterpart.  Each call enqueues
-- on the backend's stream and returns at once.
--
module Data.Array.Accelerate.Math.FFT.LLVM.HIP (

  -- * The plugin contract
  fft,
  fft1D,
  fft2D,
  fft3D,

  -- * Inverse with the 1/N fused
  fftFused,
  fft1DFused,
  fft2DFused,
  fft3DFused,

  -- * Extensions
  fft2DMany,
  fftVJP,
  fft1DVJP,
  fft2DVJP,
  fft3DVJP,
  spectralConv,
  spectralConvVJP,
  spectralConvVJPResidual,

) where

import Data.Array.Accelerate.Math.FFT.Mode
import Data.Array.Accelerate.Math.FFT.Type
import Data.Array.Accelerate.Math.FFT.LLVM.HIP.Base
import Data.Array.Accelerate.Math.FFT.LLVM.HIP.FFI

import Data.Array.Accelerate.Data.Complex
import Data.Array.Accelerate.Error
import Data.Array.Accelerate.Representation.Array
import Data.Array.Accelerate.Representation.Shape
import Data.Array.Accelerate.Sugar.Elt
import Data.Primitive.Vec

import Data.Array.Accelerate.LLVM.HIP.Foreign

import Control.Monad.Reader                                         ( asks )
import Data.Int                                                     ( Int64 )
import Foreign.C.Types                                              ( CInt )
import Foreign.Marshal.Array                                        ( withArrayLen )
import Foreign.Ptr                                                  ( Ptr )
import Prelude                                                      as P


-- The plugin contract
-- -------------------

fft :: HasCallStack
    => Mode
    -> ShapeR sh
    -> NumericR e
    -> ForeignAcc (Array sh (Vec2 e) -> Array sh (Vec2 e))
fft mode shR eR = innermost "hip.fft.many" (contractMode mode) shR eR

fft1D :: Mode -> NumericR e -> ForeignAcc (Array DIM1 (Vec2 e) -> Array DIM1 (Vec2 e))
fft1D mode eR = ForeignAcc "hip.fft1d" $ transform1D (contractMode mode) eR

fft2D :: Mode -> NumericR e -> ForeignAcc (Array DIM2 (Vec2 e) -> Array DIM2 (Vec2 e))
fft2D mode eR = ForeignAcc "hip.fft2d" $ transform2D (contractMode mode) eR

fft3D :: Mode -> NumericR e -> ForeignAcc (Array DIM3 (Vec2 e) -> Array DIM3 (Vec2 e))
fft3D mode eR = ForeignAcc "hip.fft3d" $ transform3D (contractMode mode) eR

-- The mode a plugin is asked for, as libafft's code: Reverse and Inverse both unnormalised.
contractMode :: Mode -> CInt
contractMode Forward = modeForward
contractMode Reverse = modeReverse
contractMode Inverse = modeReverse

-- The mode itself, Inverse normalised by libafft (the fused forms and the adjoints).
fullMode :: Mode -> CInt
fullMode Forward = modeForward
fullMode Reverse = modeReverse
fullMode Inverse = modeInverse


-- Inverse with the division fused into the last pass
-- --------------------------------------------------

fftFused :: HasCallStack => ShapeR sh -> NumericR e -> ForeignAcc (Array sh (Vec2 e) -> Array sh (Vec2 e))
fftFused shR eR = innermost "hip.fft.many.inverse" modeInverse shR eR

fft1DFused :: NumericR e -> ForeignAcc (Array DIM1 (Vec2 e) -> Array DIM1 (Vec2 e))
fft1DFused eR = ForeignAcc "hip.fft1d.inverse" $ transform1D modeInverse eR

fft2DFused :: NumericR e -> ForeignAcc (Array DIM2 (Vec2 e) -> Array DIM2 (Vec2 e))
fft2DFused eR = ForeignAcc "hip.fft2d.inverse" $ transform2D modeInverse eR

fft3DFused :: NumericR e -> ForeignAcc (Array DIM3 (Vec2 e) -> Array DIM3 (Vec2 e))
fft3DFused eR = ForeignAcc "hip.fft3d.inverse" $ transform3D modeInverse eR


-- Extensions
-- ----------

-- | fft2D of every matrix of a (batch, h, w) array in one call; Inverse divides by h*w.
fft2DMany :: Mode -> NumericR e -> ForeignAcc (Array DIM3 (Vec2 e) -> Array DIM3 (Vec2 e))
fft2DMany mode eR = ForeignAcc "hip.fft2d.many" $ transform dim3 eR $ \ext pin pout s ->
  case ext of
    [b, h, w] -> check "afft_fft2D_many" =<< afft_fft2D_many (fullMode mode) (dtypeCode eR) b h w pin pout s
    _         -> internalError "fft2DMany: expected a rank-3 array"

-- | Cotangent of the input of the front-end operation (@fft mode@ and so on, Inverse
-- normalised) given the cotangent of its output, in the conjugate-Wirtinger convention: the
-- conjugate-transpose transform, Forward <-> Reverse, with the 1/N of Inverse kept.
fftVJP :: HasCallStack => Mode -> ShapeR sh -> NumericR e -> ForeignAcc (Array sh (Vec2 e) -> Array sh (Vec2 e))
fftVJP mode shR eR
  | rank shR P.<= 8 = ForeignAcc "hip.fft.many.vjp" $ adjoint opFFT mode shR eR
  | otherwise       = internalError "fftVJP: only for arrays of rank 1..8"

fft1DVJP :: Mode -> NumericR e -> ForeignAcc (Array DIM1 (Vec2 e) -> Array DIM1 (Vec2 e))
fft1DVJP mode eR = ForeignAcc "hip.fft1d.vjp" $ adjoint opFFT1D mode dim1 eR

fft2DVJP :: Mode -> NumericR e -> ForeignAcc (Array DIM2 (Vec2 e) -> Array DIM2 (Vec2 e))
fft2DVJP mode eR = ForeignAcc "hip.fft2d.vjp" $ adjoint opFFT2D mode dim2 eR

fft3DVJP :: Mode -> NumericR e -> ForeignAcc (Array DIM3 (Vec2 e) -> Array DIM3 (Vec2 e))
fft3DVJP mode eR = ForeignAcc "hip.fft3d.vjp" $ adjoint opFFT3D mode dim3 eR

-- | y_b = Inverse (fft2D x_b * H) for a batch x (b, h, w) and a filter H (h, w): the row pass,
-- one fused column kernel (column transform, filter product, inverse column transform), and the
-- inverse row pass carrying 1/(h w).
spectralConv
    :: NumericR e
    -> ForeignAcc ((((), Array DIM3 (Vec2 e)), Array DIM2 (Vec2 e)) -> Array DIM3 (Vec2 e))
spectralConv eR = ForeignAcc "hip.spectral.conv" go
  where
    go (((), ax), ah) = do
      let sh@((((), b), h), w) = shape ax
      requireFilter "spectralConv" (shape ah) h w
      ay     <- allocateRemote (arrayR dim3 eR) sh
      stream <- asks hipStream
      future <- new
      liftPar $
        withArray eR ax stream $ \d_x ->
        withArray eR ah stream $ \d_h ->
        withArray eR ay stream $ \d_y ->
          liftIO $ withStreamPtr stream $ \s ->
            check "afft_spectral_conv_fwd" =<<
              afft_spectral_conv_fwd (dtypeCode eR) (i64 b) (i64 h) (i64 w)
                                     (devicePtr d_x) (devicePtr d_h) (devicePtr d_y) s
      put future ay
      return future

-- | The adjoint of 'spectralConv': given x, H and gy = dL/dy, returns
--   gx = (1/hw) F^H (conj H * F gy)  and  gH = (1/hw) sum_b conj (F x_b) * F gy_b,
-- recomputing F x instead of storing it.
spectralConvVJP
    :: NumericR e
    -> ForeignAcc (((((), Array DIM3 (Vec2 e)), Array DIM2 (Vec2 e)), Array DIM3 (Vec2 e))
                   -> (((), Array DIM3 (Vec2 e)), Array DIM2 (Vec2 e)))
spectralConvVJP eR = ForeignAcc "hip.spectral.conv.vjp" go
  where
    go ((((), ax), ah), agy) = do
      let sh@((((), b), h), w) = shape ax
          shH                  = shape ah
      requireFilter "spectralConvVJP" shH h w
      if shape agy P.== sh then return () else internalError "spectralConvVJP: gy must have the shape of x"
      agx    <- allocateRemote (arrayR dim3 eR) sh
      agH    <- allocateRemote (arrayR dim2 eR) shH
      stream <- asks hipStream
      future <- new
      liftPar $
        withArray eR ax  stream $ \d_x  ->
        withArray eR ah  stream $ \d_h  ->
        withArray eR agy stream $ \d_gy ->
        withArray eR agx stream $ \d_gx ->
        withArray eR agH stream $ \d_gH ->
          liftIO $ withStreamPtr stream $ \s ->
            check "afft_spectral_conv_vjp" =<<
              afft_spectral_conv_vjp (dtypeCode eR) (i64 b) (i64 h) (i64 w)
                                     (devicePtr d_x) (devicePtr d_h) (devicePtr d_gy)
                                     (devicePtr d_gx) (devicePtr d_gH) s
      put future (((), agx), agH)
      return future

-- | The adjoint of 'spectralConv' for the squared-error loss L = (alpha/2) sum |y - t|^2: given
-- x, H, y and the target t, the pair of 'spectralConvVJP' for gy = alpha (y - t), which the first
-- row pass forms as it loads y and t (gy is never stored).
spectralConvVJPResidual
    :: Double
    -> NumericR e
    -> ForeignAcc ((((((), Array DIM3 (Vec2 e)), Array DIM2 (Vec2 e)), Array DIM3 (Vec2 e)), Array DIM3 (Vec2 e))
                   -> (((), Array DIM3 (Vec2 e)), Array DIM2 (Vec2 e)))
spectralConvVJPResidual alpha eR = ForeignAcc "hip.spectral.conv.vjp.residual" go
  where
    go (((((), ax), ah), ay), at) = do
      let sh@((((), b), h), w) = shape ax
          shH                  = shape ah
      requireFilter "spectralConvVJPResidual" shH h w
      if shape ay P.== sh P.&& shape at P.== sh then return ()
        else internalError "spectralConvVJPResidual: y and t must have the shape of x"
      agx    <- allocateRemote (arrayR dim3 eR) sh
      agH    <- allocateRemote (arrayR dim2 eR) shH
      stream <- asks hipStream
      future <- new
      liftPar $
        withArray eR ax  stream $ \d_x  ->
        withArray eR ah  stream $ \d_h  ->
        withArray eR ay  stream $ \d_y  ->
        withArray eR at  stream $ \d_t  ->
        withArray eR agx stream $ \d_gx ->
        withArray eR agH stream $ \d_gH ->
          liftIO $ withStreamPtr stream $ \s ->
            check "afft_spectral_conv_vjp_residual" =<<
              afft_spectral_conv_vjp_residual (dtypeCode eR) (i64 b) (i64 h) (i64 w)
                                              (devicePtr d_x) 0 (devicePtr d_h) (devicePtr d_y)
                                              (devicePtr d_t) (realToFrac alpha)
                                              (devicePtr d_gx) (devicePtr d_gH) s
      put future (((), agx), agH)
      return future


-- Internals
-- ---------

-- The innermost-axis transform of any rank up to 8 (the kernel does not depend on the rank).
innermost
    :: HasCallStack
    => String
    -> CInt
    -> ShapeR sh
    -> NumericR e
    -> ForeignAcc (Array sh (Vec2 e) -> Array sh (Vec2 e))
innermost name code shR eR
  | rank shR P.<= 8 = ForeignAcc name $ transform shR eR $ \ext pin pout s ->
      withArrayLen ext $ \r pext ->
        check "afft_fft" =<< afft_fft code (dtypeCode eR) (P.fromIntegral r) pext pin pout s
  | otherwise = internalError "only for transforms of arrays of rank 1..8"

transform1D :: CInt -> NumericR e -> Array DIM1 (Vec2 e) -> Par HIP (Future (Array DIM1 (Vec2 e)))
transform1D code eR = transform dim1 eR $ \ext pin pout s ->
  case ext of
    [n] -> check "afft_fft1D" =<< afft_fft1D code (dtypeCode eR) n pin pout s
    _   -> internalError "fft1D: expected a rank-1 array"

transform2D :: CInt -> NumericR e -> Array DIM2 (Vec2 e) -> Par HIP (Future (Array DIM2 (Vec2 e)))
transform2D code eR = transform dim2 eR $ \ext pin pout s ->
  case ext of
    [h, w] -> check "afft_fft2D" =<< afft_fft2D code (dtypeCode eR) h w pin pout s
    _      -> internalError "fft2D: expected a rank-2 array"

transform3D :: CInt -> NumericR e -> Array DIM3 (Vec2 e) -> Par HIP (Future (Array DIM3 (Vec2 e)))
transform3D code eR = transform dim3 eR $ \ext pin pout s ->
  case ext of
    [d, h, w] -> check "afft_fft3D" =<< afft_fft3D code (dtypeCode eR) d h w pin pout s
    _         -> internalError "fft3D: expected a rank-3 array"

adjoint :: CInt -> Mode -> ShapeR sh -> NumericR e -> Array sh (Vec2 e) -> Par HIP (Future (Array sh (Vec2 e)))
adjoint op mode shR eR = transform shR eR $ \ext pgy pgx s ->
  withArrayLen ext $ \r pext ->
    check "afft_fft_vjp" =<< afft_fft_vjp op (fullMode mode) (dtypeCode eR) (P.fromIntegral r) pext pgy pgx s

-- Allocate a fresh output (the input stays untouched, LLVM/PTX.hs:92), lease both device
-- pointers on the backend's stream, call libafft on the raw stream with the extents listed
-- outermost first, and fill the future at once: the call is asynchronous on that stream
-- (LLVM/PTX.hs:77-106).
{-# INLINEABLE transform #-}
transform
    :: ShapeR sh
    -> NumericR e
    -> ([Int64] -> Ptr () -> Ptr () -> Ptr () -> IO ())
    -> Array sh (Vec2 e)
    -> Par HIP (Future (Array sh (Vec2 e)))
transform shR eR call ain = do
  let sh = shape ain
  aout   <- allocateRemote (arrayR shR eR) sh
  stream <- asks hipStream
  future <- new
  liftPar $
    withArray eR ain  stream $ \d_in  ->
    withArray eR aout stream $ \d_out ->
      liftIO $ withStreamPtr stream $ call (extents shR sh) (devicePtr d_in) (devicePtr d_out)
  put future aout
  return future

-- Accelerate lists a shape innermost first; libafft takes e_0 (outermost) .. e_{rank-1}.
extents :: ShapeR sh -> sh -> [Int64]
extents shR sh = P.reverse (P.map i64 (shapeToList shR sh))

arrayR :: ShapeR sh -> NumericR e -> ArrayR (Array sh (Vec2 e))
arrayR shR NumericRfloat32 = ArrayR shR (eltR @(Complex Float))
arrayR shR NumericRfloat64 = ArrayR shR (eltR @(Complex Double))

requireFilter :: HasCallStack => String -> (((), Int), Int) -> Int -> Int -> Par HIP ()
requireFilter what (((), fh), fw) h w
  | fh P.== h P.&& fw P.== w = return ()
  | otherwise                = internalError (what ++ ": the filter must have the shape of one image")

i64 :: Int -> Int64
i64 = P.fromIntegral
