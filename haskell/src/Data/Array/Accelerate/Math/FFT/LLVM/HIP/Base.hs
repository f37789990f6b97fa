{-# LANGUAGE GADTs               #-}
{-# LANGUAGE ScopedTypeVariables #-}
{-# LANGUAGE TypeApplications    #-}
{-# LANGUAGE TypeFamilies        #-}
{-# LANGUAGE TypeOperators       #-}
-- |
-- Module      : Data.Array.Accelerate.Math.FFT.LLVM.HIP.Base
-- License     : BSD3
-- Stability   : experimental
--
-- Leasing device pointers and the stream for a libafft call, as LLVM/PTX/Base.hs:32-57 does for
-- cuFFT.  The AMD-GPU backend is assumed to export the Foreign interface of accelerate-llvm-ptx
-- 1.3 (Data.Array.Accelerate.LLVM.PTX.Foreign) with HIP in place of PTX: 'withDevicePtr',
-- 'waypoint', 'Stream' (a Lifetime of the runtime stream), 'Par', 'Future', 'allocateRemote'
-- and the reader field 'hipStream'; and its HIP binding to mirror the cuda package's
-- Foreign.CUDA.Ptr ('DevicePtr', 'useDevicePtr') and stream type ('useStream').
--
module Data.Array.Accelerate.Math.FFT.LLVM.HIP.Base
  where

import Data.Array.Accelerate.Math.FFT.Type
import Data.Array.Accelerate.Math.FFT.LLVM.HIP.FFI                  ( afft_strerror )

import Data.Array.Accelerate.Array.Data
import Data.Array.Accelerate.Error
import Data.Array.Accelerate.Lifetime
import Data.Array.Accelerate.Representation.Array
import Data.Array.Accelerate.Type
import Data.Primitive.Vec

import Data.Array.Accelerate.LLVM.HIP.Foreign

import Foreign.HIP.Ptr                                              ( DevicePtr, useDevicePtr )
import qualified Foreign.HIP.Stream                                 as HIP

import Foreign.C.String                                             ( peekCString )
import Foreign.C.Types                                              ( CInt )
import Foreign.Ptr                                                  ( Ptr, castPtr )


-- | Lease the device pointer of an array of complex numbers for a computation on stream @s@.
-- When the computation has been enqueued, an event is recorded on @s@ (the waypoint): the
-- memory manager keeps the buffer alive until the device has passed it.
{-# INLINE withArray #-}
withArray
    :: NumericR e
    -> Array sh (Vec2 e)
    -> Stream
    -> (DevicePtr e -> LLVM HIP b)
    -> LLVM HIP b
withArray eR (Array _ adata) = withArrayData eR adata

{-# INLINE withArrayData #-}
withArrayData
    :: NumericR e
    -> ArrayData (Vec2 e)
    -> Stream
    -> (DevicePtr e -> LLVM HIP b)
    -> LLVM HIP b
withArrayData NumericRfloat32 ad s k =
  withDevicePtr (singleType @Float) ad $ \p -> do
    r <- k p
    e <- waypoint s
    return (Just e, r)
withArrayData NumericRfloat64 ad s k =
  withDevicePtr (singleType @Double) ad $ \p -> do
    r <- k p
    e <- waypoint s
    return (Just e, r)

-- | The raw device address of a leased buffer, as libafft takes it (an interleaved Vec2 array).
{-# INLINE devicePtr #-}
devicePtr :: DevicePtr e -> Ptr ()
devicePtr = castPtr . useDevicePtr

-- | The raw hipStream_t of the backend's stream, alive for the duration of the action.  libafft
-- keeps no stream between calls: each call enqueues on the stream it is given.
{-# INLINE withStreamPtr #-}
withStreamPtr :: Stream -> (Ptr () -> IO a) -> IO a
withStreamPtr s k = withLifetime s (k . HIP.useStream)

-- | A nonzero libafft status becomes an internal error naming the call, as the reference plugins
-- report an unsupported rank (LLVM/PTX.hs:61).
check :: HasCallStack => String -> CInt -> IO ()
check _    0  = return ()
check what st = do
  msg <- peekCString =<< afft_strerror st
  internalError (what ++ ": " ++ msg ++ " (libafft status " ++ show st ++ ")")

-- | Element-type code of libafft (AFFT_C64 / AFFT_C128).
dtypeCode :: NumericR e -> CInt
dtypeCode NumericRfloat32 = 0
dtypeCode NumericRfloat64 = 1
