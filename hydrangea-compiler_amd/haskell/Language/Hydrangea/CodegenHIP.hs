{-# LANGUAGE OverloadedStrings #-}
-- |
-- Module: Language.Hydrangea.CodegenHIP
--
-- The HIP offload for the C backend.  A CFG-to-CFG pass replaces each loop of
-- one of three shapes with a single runtime call into libhydhip (MI355X
-- kernels).  The C backend then prints the program as usual: it prints an
-- 'RCall' as a plain C call (src/Language/Hydrangea/CodegenC.hs:2340-2344),
-- and include/hyd_hip_glue.h defines the called helpers.  Every other loop
-- stays the C backend's OpenMP loop, so the result is always a complete
-- program.
--
-- Shapes (single iterator from 0, bound a variable or a constant):
--
-- * scatter-add, the loop CodegenC.hs:1576-1611 detects, unguarded, whose
--   index and value are loads at the iterator.  The parallelizer tagged it
--   'ParallelScatterAtomicAddInt' / 'ParallelScatterPrivatizedIntAdd'
--   (int64, 'CAdd') or the Float variants (fp64, 'CAddF'), CFG.hs:184-191.
--   It becomes @_discarded = hydg_scatter_add_{i64,f64}(dst, idx, vals, n)@.
-- * the fused weighted histogram, the same scatter loop (int64) whose body
--   loads @x@ at the iterator, routes it by @x % nbins@ ('CMod') and adds
--   @x * x@ ('CMul').  It becomes
--   @_discarded = hydg_weighted_histogram_i64(dst, data, nbins, n)@.
-- * dot: a reduction loop ('lsRed' with 'RAdd') whose body loads @x@ and @y@
--   at the iterator and adds @x *. y@ to the accumulator.  It becomes
--   @acc = hydg_dot_f64(xs, ys, n, acc)@.
--
-- UNTESTED: GHC is not available where this module was written.  It follows
-- the interfaces of CFG.hs (LoopSpec :251-259, ExecPolicy :239-244,
-- ParallelSpec :199-204, Stmt :262-270, rewriteStmtsWith :278-288),
-- CFGCore.hs (RHS :140-169, BinOp :33-74, Atom :77-85, Redop :236),
-- CFGPipeline.hs (metalPipelineWithTiling :224-226) and CodegenC.hs
-- (codegenProgramWithOptions :161-162, CodegenArtifacts :52-56).  The C side
-- of every call it emits is tested from a gcc-built host.
module Language.Hydrangea.CodegenHIP
  ( HipSite (..)
  , hipPipeline
  , hipOffloadProgram
  , hipOffloadSites
  , codegenHIPWithOptions
  ) where

import Control.Applicative ((<|>))
import Control.Monad (guard)
import Data.ByteString (ByteString)
import qualified Data.Map.Strict as M
import Language.Hydrangea.CFG
import Language.Hydrangea.CFGCore
import Language.Hydrangea.CFGPipeline (metalPipelineWithTiling)
import Language.Hydrangea.CodegenC (CodegenArtifacts (..), CodegenOptions, codegenProgramWithOptions)

-- | One offloaded loop: the helper it calls and the arrays it passes.
data HipSite = HipSite
  { hsHelper :: ByteString
  , hsArrays :: [CVar]
  }
  deriving (Eq, Show)

-- | Optimize and parallelize without SIMD vectorization (the GPU supplies
-- the lanes, and vector loops would hide the shapes), then offload.
hipPipeline :: Program -> Program
hipPipeline = hipOffloadProgram . metalPipelineWithTiling False

hipOffloadProgram :: Program -> Program
hipOffloadProgram (Program procs) =
  Program [p {procBody = rewriteStmtsWith () id (const rewrite) (procBody p)} | p <- procs]
  where
    rewrite s = maybe [s] snd (offload s)

-- | The loops 'hipOffloadProgram' replaces, in program order.
hipOffloadSites :: Program -> [HipSite]
hipOffloadSites (Program procs) = concatMap (goStmts . procBody) procs
  where
    goStmts = concatMap goStmt
    goStmt s = case offload s of
      Just (site, _) -> [site]
      Nothing -> case s of
        SLoop _ body -> goStmts body
        SParallelRegion body -> goStmts body
        SIf _ t e -> goStmts t ++ goStmts e
        _ -> []

-- | The C backend's output for the offloaded program, with the helpers'
-- header included.
codegenHIPWithOptions :: CodegenOptions -> Program -> Either String CodegenArtifacts
codegenHIPWithOptions opts prog = do
  arts <- codegenProgramWithOptions opts (hipOffloadProgram prog)
  pure arts {codegenSource = "#include \"hyd_hip_glue.h\"\n" ++ codegenSource arts}

offload :: Stmt -> Maybe (HipSite, [Stmt])
offload (SLoop spec body) = do
  [i] <- Just (lsIters spec)
  [bound] <- Just (lsBounds spec)
  guard (all (== IConst 0) (lsOrigins spec))
  n <- boundAtom bound
  scatterAdd spec i n body <|> weightedHistogram spec i n body <|> dot spec i n body
offload _ = Nothing

boundAtom :: IndexExpr -> Maybe Atom
boundAtom (IVar v) = Just (AVar v)
boundAtom (IConst k) = Just (AInt k)
boundAtom _ = Nothing

-- | Variables bound to a load at the iterator, and the array each reads.
loadsAt :: CVar -> [Stmt] -> Maybe (M.Map CVar CVar)
loadsAt i = fmap M.fromList . mapM one
  where
    one (SAssign v (RArrayLoad (AVar arr) (AVar j))) | j == i = Just (v, arr)
    one _ = Nothing

scatterAdd :: LoopSpec -> CVar -> Atom -> [Stmt] -> Maybe (HipSite, [Stmt])
scatterAdd spec i n body = do
  Parallel ps <- Just (lsExec spec)
  (op, helper) <- strategyHelper (psStrategy ps)
  [l1, l2, SAssign old (RArrayLoad (AVar dst) (AVar ix)), SAssign new (RBinOp op' a b), SArrayWrite (AVar dst') (AVar ix') (AVar new')] <-
    Just body
  loads <- loadsAt i [l1, l2]
  guard (op' == op && dst' == dst && ix' == ix && new' == new)
  v <- case (a, b) of
    (AVar x, AVar y) | x == old -> Just y
                     | y == old -> Just x
    _ -> Nothing
  idxArr <- M.lookup ix loads
  valArr <- M.lookup v loads
  guard (ix /= v && dst /= idxArr && dst /= valArr)
  let call = RCall helper [AVar dst, AVar idxArr, AVar valArr, n]
  pure (HipSite helper [dst, idxArr, valArr], [SAssign "_discarded" call])
  where
    strategyHelper s = case s of
      ParallelScatterAtomicAddInt -> Just (CAdd, "hydg_scatter_add_i64")
      ParallelScatterPrivatizedIntAdd -> Just (CAdd, "hydg_scatter_add_i64")
      ParallelScatterAtomicAddFloat -> Just (CAddF, "hydg_scatter_add_f64")
      ParallelScatterPrivatizedFloatAdd -> Just (CAddF, "hydg_scatter_add_f64")
      _ -> Nothing

weightedHistogram :: LoopSpec -> CVar -> Atom -> [Stmt] -> Maybe (HipSite, [Stmt])
weightedHistogram spec i n body = do
  Parallel ps <- Just (lsExec spec)
  guard (psStrategy ps `elem` [ParallelScatterAtomicAddInt, ParallelScatterPrivatizedIntAdd])
  [s1, s2, s3, SAssign old (RArrayLoad (AVar dst) (AVar r)), SAssign new (RBinOp CAdd a b), SArrayWrite (AVar dst') (AVar r') (AVar new')] <-
    Just body
  guard (dst' == dst && r' == r && new' == new)
  w <- case (a, b) of
    (AVar p, AVar q) | p == old -> Just q
                     | q == old -> Just p
    _ -> Nothing
  -- the load of x at i, the route r = x % nbins and the weight w = x * x,
  -- the last two in either order
  (x, dataArr) <- case s1 of
    SAssign v (RArrayLoad (AVar arr) (AVar j)) | j == i -> Just (v, arr)
    _ -> Nothing
  let route (SAssign v (RBinOp CMod (AVar x') nb)) | v == r && x' == x = Just nb
      route _ = Nothing
      weight (SAssign v (RBinOp CMul (AVar p) (AVar q))) = v == w && p == x && q == x
      weight _ = False
  nbins <- (route s2 <* guard (weight s3)) <|> (route s3 <* guard (weight s2))
  guard (dst /= dataArr && r /= w)
  let helper = "hydg_weighted_histogram_i64"
  pure (HipSite helper [dst, dataArr], [SAssign "_discarded" (RCall helper [AVar dst, AVar dataArr, nbins, n])])

dot :: LoopSpec -> CVar -> Atom -> [Stmt] -> Maybe (HipSite, [Stmt])
dot spec i n body = do
  rs <- lsRed spec
  guard (rsRedop rs == RAdd)
  case lsExec spec of
    Vector _ -> Nothing
    _ -> Just ()
  let acc = rsAccVar rs
  [l1, l2, SAssign p (RBinOp CMulF (AVar x) (AVar y)), SAssign acc' (RBinOp CAddF a b)] <- Just body
  guard (acc' == acc && (a, b) `elem` [(AVar acc, AVar p), (AVar p, AVar acc)])
  loads <- loadsAt i [l1, l2]
  xs <- M.lookup x loads
  ys <- M.lookup y loads
  guard (x /= y && p /= acc)
  let helper = "hydg_dot_f64"
  pure (HipSite helper [xs, ys], [SAssign acc (RCall helper [AVar xs, AVar ys, n, AVar acc])])
