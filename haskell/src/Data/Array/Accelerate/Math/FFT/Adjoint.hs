{-# LANGUAGE CPP                 #-}
{-# LANGUAGE FlexibleContexts    #-}
{-# LANGUAGE ScopedTypeVariables #-}
{-# LANGUAGE TypeApplications    #-}
{-# LANGUAGE TypeOperators       #-}
-- |
-- Module      : Data.Array.Accelerate.Math.FFT.Adjoint
-- License     : BSD3
-- Stability   : experimental
--
-- Adjoints (vector-Jacobian products) of the transforms of Data.Array.Accelerate.Math.FFT, and a
-- differentiable 2D spectral convolution over a batch of images.  Each function is a chain of
-- foreign layers over a pure Accelerate definition, as the transforms themselves are
-- (FFT.hs:73-80): the HIP backend runs libafft's fused kernels, every other backend evaluates the
-- pure definition, which uses the transforms of Data.Array.Accelerate.Math.FFT (and so their own
-- FFTW and cuFFT plugins).
--
-- Convention: for y = op x, the cotangent of x given the cotangent gy of y is op^H gy
-- (conjugate-Wirtinger, as torch autograd).  The conjugate transpose of Forward is Reverse and the
-- other way round; Inverse is Reverse / N, so its adjoint is Forward / N.
--
module Data.Array.Accelerate.Math.FFT.Adjoint (

  fftVJP,
  fft1DVJP,
  fft2DVJP,
  fft3DVJP,

  spectralConv,
  spectralConvVJP,
  spectralConvVJPResidual,

) where

import Data.Array.Accelerate                                        as A
import Data.Array.Accelerate.Data.Complex
import Data.Array.Accelerate.Math.FFT
import Data.Array.Accelerate.Math.FFT.Type                          ( numericR )
import qualified Data.Array.Accelerate.Sugar.Shape                  as A ( rank, shapeR )

#ifdef ACCELERATE_LLVM_HIP_BACKEND
import qualified Data.Array.Accelerate.Math.FFT.LLVM.HIP            as HIP
#endif

import Prelude                                                      as P


-- | Cotangent of the input of @'fft' mode@ (innermost axis).
fftVJP :: forall sh e. (Shape sh, Slice sh, Numeric e)
       => Mode
       -> Acc (Array (sh:.Int) (Complex e))
       -> Acc (Array (sh:.Int) (Complex e))
fftVJP mode gy
  = let
        scale = A.fromIntegral (indexHead (shape gy))
        rank  = A.rank @(sh:.Int)
        shR   = A.shapeR @(sh:.Int)
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  (if rank P.<= 8 then foreignAcc (HIP.fftVJP mode shR eR) else id) $
#endif
                  adjointBy fft scale mode
    in
    go gy

-- | Cotangent of the input of @'fft1D' mode@.
fft1DVJP :: forall e. Numeric e
         => Mode
         -> Acc (Array DIM1 (Complex e))
         -> Acc (Array DIM1 (Complex e))
fft1DVJP mode gy
  = let
        scale = A.fromIntegral (A.length gy)
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.fft1DVJP mode eR) $
#endif
                  adjointBy fft1D scale mode
    in
    go gy

-- | Cotangent of the input of @'fft2D' mode@.
fft2DVJP :: forall e. Numeric e
         => Mode
         -> Acc (Array DIM2 (Complex e))
         -> Acc (Array DIM2 (Complex e))
fft2DVJP mode gy
  = let
        scale = A.fromIntegral (A.size gy)
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.fft2DVJP mode eR) $
#endif
                  adjointBy fft2D scale mode
    in
    go gy

-- | Cotangent of the input of @'fft3D' mode@.
fft3DVJP :: forall e. Numeric e
         => Mode
         -> Acc (Array DIM3 (Complex e))
         -> Acc (Array DIM3 (Complex e))
fft3DVJP mode gy
  = let
        scale = A.fromIntegral (A.size gy)
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.fft3DVJP mode eR) $
#endif
                  adjointBy fft3D scale mode
    in
    go gy

-- The adjoint of a front-end transform, from the transform itself.
adjointBy :: (Shape sh, Numeric e)
          => (Mode -> Acc (Array sh (Complex e)) -> Acc (Array sh (Complex e)))
          -> Exp (Complex e)
          -> Mode
          -> Acc (Array sh (Complex e))
          -> Acc (Array sh (Complex e))
adjointBy op scale mode gy =
  case mode of
    Forward -> op Reverse gy
    Reverse -> op Forward gy
    Inverse -> A.map (/scale) (op Forward gy)


-- | y_b = Inverse (fft2D x_b * H) for every image x_b of a (b, h, w) batch and a filter H of
-- shape (h, w).
spectralConv :: forall e. Numeric e
             => Acc (Array DIM3 (Complex e))
             -> Acc (Array DIM2 (Complex e))
             -> Acc (Array DIM3 (Complex e))
spectralConv x h
  = let
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.spectralConv eR) $
#endif
                  \xh -> convolve (afst xh) (asnd xh)
    in
    go (A.lift (x, h))

-- | The adjoint of 'spectralConv': given x, H and gy = dL/dy, the pair
--   ( gx = (1/hw) F^H (conj H * F gy),  gH = (1/hw) sum_b conj (F x_b) * F gy_b ).
-- The HIP backend recomputes F x inside its fused kernels instead of storing it.
spectralConvVJP :: forall e. Numeric e
                => Acc (Array DIM3 (Complex e))
                -> Acc (Array DIM2 (Complex e))
                -> Acc (Array DIM3 (Complex e))
                -> Acc (Array DIM3 (Complex e), Array DIM2 (Complex e))
spectralConvVJP x h gy
  = let
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.spectralConvVJP eR) $
#endif
                  \t -> let (x', h', g') = A.unlift t
                                             :: ( Acc (Array DIM3 (Complex e))
                                                , Acc (Array DIM2 (Complex e))
                                                , Acc (Array DIM3 (Complex e)) )
                        in  convolveVJP x' h' g'
    in
    go (A.lift (x, h, gy))

-- | 'spectralConvVJP' for the squared-error loss L = (alpha/2) sum |y - t|^2, from y = spectralConv
-- x H and the target t: the gradient alpha (y - t) is formed by the HIP backend as its first row
-- pass loads y and t, and never stored.
spectralConvVJPResidual :: forall e. Numeric e
                        => Double
                        -> Acc (Array DIM3 (Complex e))
                        -> Acc (Array DIM2 (Complex e))
                        -> Acc (Array DIM3 (Complex e))
                        -> Acc (Array DIM3 (Complex e))
                        -> Acc (Array DIM3 (Complex e), Array DIM2 (Complex e))
spectralConvVJPResidual alpha x h y t
  = let
        eR    = numericR @e
        go    =
#ifdef ACCELERATE_LLVM_HIP_BACKEND
                  foreignAcc (HIP.spectralConvVJPResidual alpha eR) $
#endif
                  \a -> let (x', h', y', t') = A.unlift a
                                                :: ( Acc (Array DIM3 (Complex e))
                                                   , Acc (Array DIM2 (Complex e))
                                                   , Acc (Array DIM3 (Complex e))
                                                   , Acc (Array DIM3 (Complex e)) )
                             k = A.constant (P.realToFrac alpha :+ 0)
                         in  convolveVJP x' h' (A.zipWith (\u v -> k * (u - v)) y' t')
    in
    go (A.lift (x, h, y, t))


-- Pure definitions
-- ----------------

convolve :: Numeric e
         => Acc (Array DIM3 (Complex e))
         -> Acc (Array DIM2 (Complex e))
         -> Acc (Array DIM3 (Complex e))
convolve x h =
  let Z :. b :. _ :. _ = unlift (shape x) :: Z :. Exp Int :. Exp Int :. Exp Int
      hb               = A.replicate (A.lift (Z :. b :. All :. All)) h
  in
  fft2DBatch Inverse (A.zipWith (*) (fft2DBatch Forward x) hb)

convolveVJP :: Numeric e
            => Acc (Array DIM3 (Complex e))
            -> Acc (Array DIM2 (Complex e))
            -> Acc (Array DIM3 (Complex e))
            -> Acc (Array DIM3 (Complex e), Array DIM2 (Complex e))
convolveVJP x h gy =
  let Z :. b :. n :. m = unlift (shape x) :: Z :. Exp Int :. Exp Int :. Exp Int
      scale            = A.fromIntegral (n * m)
      hb               = A.replicate (A.lift (Z :. b :. All :. All)) h
      fg               = fft2DBatch Forward gy
      fx               = fft2DBatch Forward x
      gx               = fft2DBatch Inverse (A.zipWith (*) (A.map conjugate hb) fg)
      gH               = A.map (/scale) (A.sum (batchInnermost (A.zipWith (*) (A.map conjugate fx) fg)))
  in
  A.lift (gx, gH)

-- fft2D of every matrix of a (b, h, w) array: the innermost-axis transform, then the same over
-- the columns as rows of the transposed matrices; Inverse divides by w and then by h.
fft2DBatch :: Numeric e
           => Mode
           -> Acc (Array DIM3 (Complex e))
           -> Acc (Array DIM3 (Complex e))
fft2DBatch mode = swapInner . fft mode . swapInner . fft mode

-- (b, i, j) -> (b, j, i); its own inverse
swapInner :: Elt a => Acc (Array DIM3 a) -> Acc (Array DIM3 a)
swapInner arr = backpermute (swapIx (shape arr)) swapIx arr
  where
    swapIx :: Exp DIM3 -> Exp DIM3
    swapIx ix =
      let Z :. b :. i :. j = unlift ix :: Z :. Exp Int :. Exp Int :. Exp Int
      in  index3 b j i

-- (b, i, j) -> (i, j, b): the batch axis innermost, for a sum over it
batchInnermost :: Elt a => Acc (Array DIM3 a) -> Acc (Array DIM3 a)
batchInnermost arr = backpermute sh from arr
  where
    sh :: Exp DIM3
    sh =
      let Z :. b :. i :. j = unlift (shape arr) :: Z :. Exp Int :. Exp Int :. Exp Int
      in  index3 i j b
    --
    from :: Exp DIM3 -> Exp DIM3
    from ix =
      let Z :. i :. j :. b = unlift ix :: Z :. Exp Int :. Exp Int :. Exp Int
      in  index3 b i j
