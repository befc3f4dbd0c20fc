{-# LANGUAGE OverloadedStrings #-}
-- |
-- Module: Language.Hydrangea.CodegenHIP
--
-- The MI355X backend beside CodegenC and CodegenMSL.  After the Metal-style
-- pipeline (optimize + parallelize without SIMD vectors, CFGPipeline.hs:
-- 224-226) an offload pass picks, in every procedure, the outermost loops
-- whose iterations run independently and whose bodies the device can run
-- exactly as the printed C does; each becomes one call of a generated
-- launcher (an 'RCall', printed by CodegenC.hs:2340-2344 as a plain C call).
-- The loops themselves go out as 'HipSite' records, one `show` per line: the
-- kernel generator (hydrangea-compiler_amd/hydgen) reads that text, runs the
-- GPU-schedule pass (grid, workgroup, elements per thread, reduction tree,
-- scatter merge) and prints the HIP kernels and launchers.  Every other
-- statement stays CodegenC's host C.
--
-- Offloaded (the kernel loop of a site):
--
-- * a loop with a ReductionSpec (CFG.hs:207-212), or a LoopReductionWrapper
--   of trip count 1 around one (Lowering.hs:1150-1182): the site returns the
--   accumulator, which must be assigned before the loop;
-- * a loop the parallelizer marked 'Parallel' (CFG.hs:239-244) or whose role
--   is 'LoopMap' / 'LoopMapReduction' (independent iterations by
--   construction); a combining scatter strategy only with the (+) / (+.)
--   combine of Lowering.hs:3553-3557;
-- and only when every statement of its body is one the device runs as C
-- does (three-address arithmetic, loads and stores, ifs, inner loops; no
-- log/sin/cos/erf, whose glibc results a device has no exact form of; no
-- calls, allocation, records or pairs), when no variable it assigns is read
-- after it (the accumulator aside), and when it is not inside a parallel host
-- loop or region.
--
-- An iterate (Lowering.hs:840-857: a serial LoopIterate loop whose step
-- allocates the new state once, evaluates scalars of the state's shape,
-- branches on them, runs loops offloadable as above and ends with
-- cur = next) is offloaded whole, as one site returning the final state:
-- hydgen runs a split stencil's steps several per launch over LDS tiles
-- (temporal blocking), and any other step's loops in order.
--
-- Adjacent calls of map sites (loops without a result, nothing between
-- them: the region loops of a split stencil) become one call of a group
-- launcher, listed as a 'HipGroup' line; hydgen runs the members in one
-- launch when it finds their accesses independent at run time.
--
-- Residency: arrays stay in HBM between sites.  The pass brings an array's
-- host copy up to date (hyd_array_sync_host) before a host statement reads or
-- writes its elements and before it is returned, marks host writes
-- (hyd_array_mark_host_dirty), marks array parameters on entry
-- (hyd_array_enter_host: the caller may have rewritten them), and allocates
-- the 8-byte-element arrays a site touches with hyd_array_alloc_device8, so
-- that a kernel writing a fresh array uploads nothing.  The runtime calls are
-- no-ops when nothing is out of date.
--
-- UNTESTED here: GHC is not available where this module was written.  It
-- follows CFG.hs (Stmt :262-270, LoopSpec :251-259, rewriteStmtsWith
-- :278-288), CFGCore.hs (RHS :140-169, Atom :77-85, CType :103-113),
-- CFGTyping.hs (recoverProcTypeEnv :363, inferProgramReturnTypes :341,
-- buildCallParamTypes :354) and CodegenC.hs (codegenProgramWithOptions :161,
-- CodegenArtifacts :52-56).  Everything downstream of the site text -
-- hydgen, the generated kernels, the compile driver - is tested against the
-- C oracle on the GPU (tests/test_gpu_hydgen.py, tests/test_gpu_hydhip_cc.py).
module Language.Hydrangea.CodegenHIP
  ( HipSite (..)
  , HipGroup (..)
  , HIPArtifacts (..)
  , hipPipeline
  , hipOffloadProgram
  , offloadableLoop
  , offloadableIterate
  , codegenHIPWithOptions
  , renderSites
  ) where

import Control.Monad.State.Strict (State, get, put, runState)
import qualified Data.ByteString.Lazy.Char8 as BS
import Data.List (nub, sort)
import qualified Data.Map.Strict as M
import Data.Maybe (isJust, isNothing, mapMaybe)
import qualified Data.Set as S
import Language.Hydrangea.CFG
import Language.Hydrangea.CFGCore
import Language.Hydrangea.CFGPipeline (metalPipelineWithTiling)
import Language.Hydrangea.CFGTyping (TypeEnv, buildCallParamTypes, inferProgramReturnTypes, recoverProcTypeEnv)
import Language.Hydrangea.CodegenC (CodegenArtifacts (..), CodegenOptions, codegenProgramWithOptions)

-- | A run of adjacent map sites the printed C calls through one launcher
-- (hydgen/cfg.cpp read_groups): hydgen runs them in one launch when their
-- accesses are independent at run time - the split stencil's region loops
-- are the common case - and in order otherwise.
data HipGroup = HipGroup
  { hgName :: CVar
  , hgSites :: [CVar]
  }
  deriving (Eq, Show)

-- | One offloaded loop, as hydgen reads it (hydgen/cfg.cpp).
data HipSite = HipSite
  { hsName :: CVar
  , hsArgs :: [(CVar, CType)]
  , hsResult :: Maybe (CVar, CType)
  , hsTypes :: M.Map CVar CType
  , hsLoop :: Stmt
  }
  deriving (Eq, Show)

data HIPArtifacts = HIPArtifacts
  { hipC :: String           -- ^ the printed C, launchers called by name
  , hipSites :: [HipSite]    -- ^ the offloaded loops
  , hipGroups :: [HipGroup]  -- ^ runs of adjacent map sites called as one
  , hipHeader :: Maybe String
  }

-- | Optimize and parallelize without SIMD vectors: the GPU supplies the
-- lanes, and vector loops would hide the loop shapes.
hipPipeline :: Program -> Program
hipPipeline = metalPipelineWithTiling False

renderSites :: [HipSite] -> [HipGroup] -> String
renderSites sites groups = unlines (map show sites ++ map show groups)

-- | The C backend's output for the offloaded program; the printed C includes
-- the residency runtime, and the launchers' declarations come from hydgen
-- (--header), force-included by the compile driver.
codegenHIPWithOptions :: CodegenOptions -> Program -> Either String HIPArtifacts
codegenHIPWithOptions opts prog = do
  let (prog', sites, groups) = hipOffloadProgram prog
  arts <- codegenProgramWithOptions opts prog'
  pure HIPArtifacts
    { hipC = "#include \"hyd_hip_runtime.h\"\n" ++ codegenSource arts
    , hipSites = sites
    , hipGroups = groups
    , hipHeader = codegenHeader arts
    }

-- ---------------------------------------------------------------------------
-- The pass

-- | Once any loop of the program is offloaded, every procedure gets the
-- residency calls: a procedure without sites may still read or write, on
-- the host, arrays that another procedure's kernels left in HBM.
hipOffloadProgram :: Program -> (Program, [HipSite], [HipGroup])
hipOffloadProgram prog@(Program procs) =
  let callTypes = inferProgramReturnTypes prog
      callParams = buildCallParamTypes callTypes procs
      results = map (offloadProc callTypes callParams) procs
      sites = concat [ss | (_, ss, _) <- results]
      groups = concat [gs | (_, _, gs) <- results]
   in if null sites then (prog, [], []) else (Program [q | (q, _, _) <- results], sites, groups)

discarded :: CVar
discarded = "_discarded"

callStmt :: CVar -> [Atom] -> Stmt
callStmt fn args = SAssign discarded (RCall fn args)

offloadProc :: M.Map CVar CType -> M.Map CVar [Maybe CType] -> Proc -> (Proc, [HipSite], [HipGroup])
offloadProc callTypes callParams p =
  let env = recoverProcTypeEnv callTypes callParams p
      isArr v = case M.lookup v env of
        Just (CTArray _) -> True
        _ -> False
      body = procBody p
      (body', (_, sites, groups)) = runState (goBlock env isArr (procName p) body body) (0 :: Int, [], [])
      siteArrays = S.fromList [v | s <- sites, (v, CTArray _) <- hsArgs s]
      entry = [callStmt "hyd_array_enter_host" [AVar v] | v <- procParams p, isArr v]
      body'' = entry ++ map (deviceAllocs env siteArrays) body'
      env' = M.insert discarded CTInt64 (procTypeEnv p)
   in (p {procBody = body'', procTypeEnv = env'}, reverse sites, reverse groups)

-- | Fresh 8-byte-element arrays that a site touches are allocated in HBM.
deviceAllocs :: TypeEnv -> S.Set CVar -> Stmt -> Stmt
deviceAllocs env siteArrays = go
  where
    go s = case s of
      SAssign v (RArrayAlloc shp)
        | v `S.member` siteArrays, Just (CTArray t) <- M.lookup v env, t `elem` [CTInt64, CTDouble, CTBool] ->
            SAssign v (RCall "hyd_array_alloc_device8" [shp])
      SLoop spec b -> SLoop spec (map go b)
      SIf c t e -> SIf c (map go t) (map go e)
      SParallelRegion b -> SParallelRegion (map go b)
      _ -> s

type Pass = State (Int, [HipSite], [HipGroup])

-- | Statements of one block of procedure `pname`.  `whole` is the
-- procedure body (for the live-out check).
goBlock :: TypeEnv -> (CVar -> Bool) -> CVar -> [Stmt] -> [Stmt] -> Pass [Stmt]
goBlock env isArr pname whole stmts = mapM one stmts >>= groupRuns . concat
  where
    one s = case offloadableIterate env whole s `orElse` offloadableLoop env whole s of
      Just mk -> do
        (k, sites, groups) <- get
        let site = mk (siteName k)
        put (k + 1, site : sites, groups)
        pure [siteCall site]
      Nothing -> case s of
        -- host loops and branches that hold sites: recurse (the iterate
        -- loop around its step is the common case); a parallel host loop or
        -- region is never entered, so no launch runs on several threads
        SLoop spec b | lsExec spec == Serial, containsSite b -> do
          b' <- goBlock env isArr pname whole b
          pure [SLoop spec b']
        SIf c t e | containsSite t || containsSite e -> do
          t' <- goBlock env isArr pname whole t
          e' <- goBlock env isArr pname whole e
          pure [SIf c t' e']
        SReturn a -> pure (syncReturned a ++ [s])
        _ -> pure (hostStatement isArr s)
    containsSite = any (\t -> isJust (offloadableLoop env whole t) || isJust (offloadableIterate env whole t) || nested t)
    orElse a b = maybe b Just a
    nested t = case t of
      SLoop spec b | lsExec spec == Serial -> containsSite b
      SIf _ a b -> containsSite a || containsSite b
      _ -> False
    siteName k = BS.pack "hydk_" <> sanitizeName pname <> BS.pack ("_" ++ show k)
    syncReturned a = case a of
      AVar v | isArr v -> [callStmt "hyd_array_sync_host" [AVar v]]
      AVar v -> [callStmt "hyd_array_sync_host" [AVar f] | f <- recordArrays v]
      _ -> []
    recordArrays v = nub (concat [ [x | AVar x <- map snd fs, isArr x] | SAssign w (RRecord fs) <- flatten whole, w == v ]
                          ++ concat [ [x | AVar x <- [a, b], isArr x] | SAssign w (RPairMake _ _ a b) <- flatten whole, w == v ])

-- | Runs of two or more adjacent calls of map sites (no result, nothing
-- between them) become one call of a group launcher whose arguments are the
-- members' arguments, each once, in order of first appearance - the order
-- hydgen gives the group launcher's parameters.
groupRuns :: [Stmt] -> Pass [Stmt]
groupRuns stmts = do
  (_, sites, _) <- get
  let mapSites = S.fromList [hsName s | s <- sites, isNothing (hsResult s)]
      isMapCall t = case t of
        SAssign v (RCall fn _) -> v == discarded && fn `S.member` mapSites
        _ -> False
      go [] = pure []
      go xs@(y : ys)
        | isMapCall y =
            let (run, rest) = span isMapCall xs
             in case run of
                  (_ : _ : _) -> do
                    (k, ss, gs) <- get
                    let names = [fn | SAssign _ (RCall fn _) <- run]
                        gname = head names <> BS.pack "_group"
                        args = nub (concat [as | SAssign _ (RCall _ as) <- run])
                    put (k, ss, HipGroup gname names : gs)
                    rest' <- go rest
                    pure (callStmt gname args : rest')
                  _ -> (run ++) <$> go rest
        | otherwise = (y :) <$> go ys
  go stmts

sanitizeName :: CVar -> CVar
sanitizeName = BS.map (\c -> if c `elem` (['a' .. 'z'] ++ ['A' .. 'Z'] ++ ['0' .. '9'] ++ "_") then c else '_')

-- | A host statement with the residency calls around it: host copies
-- brought up to date before it touches array elements or hands arrays to a
-- procedure of the program (the library's own hyd_* functions bring device
-- copies back themselves), host writes marked after it.
hostStatement :: (CVar -> Bool) -> Stmt -> [Stmt]
hostStatement isArr s =
  let touched = sort (nub [v | v <- elementAccesses s ++ procCallArrays s, isArr v])
      written = sort (nub [v | v <- elementWrites s, isArr v])
   in [callStmt "hyd_array_sync_host" [AVar v] | v <- touched]
        ++ [s]
        ++ [callStmt "hyd_array_mark_host_dirty" [AVar v] | v <- written]

elementAccesses :: Stmt -> [CVar]
elementAccesses s = [v | t <- flatten [s], v <- acc t]
  where
    acc t = case t of
      SAssign _ (RArrayLoad (AVar a) _) -> [a]
      SAssign _ (RVecLoad (AVar a) _) -> [a]
      SAssign _ (RVecStore (AVar a) _ _) -> [a]
      SArrayWrite (AVar a) _ _ -> [a]
      _ -> []

procCallArrays :: Stmt -> [CVar]
procCallArrays s = [v | SAssign _ (RCall fn as) <- flatten [s], not ("hyd" `BS.isPrefixOf` fn), AVar v <- as]

elementWrites :: Stmt -> [CVar]
elementWrites s = [v | t <- flatten [s], v <- wr t]
  where
    wr t = case t of
      SArrayWrite (AVar a) _ _ -> [a]
      SAssign _ (RVecStore (AVar a) _ _) -> [a]
      _ -> []

flatten :: [Stmt] -> [Stmt]
flatten = concatMap f
  where
    f s = s : case s of
      SLoop _ b -> flatten b
      SIf _ t e -> flatten t ++ flatten e
      SParallelRegion b -> flatten b
      _ -> []

siteCall :: HipSite -> Stmt
siteCall site =
  let args = map (AVar . fst) (hsArgs site)
   in case hsResult site of
        Just (acc, _) -> SAssign acc (RCall (hsName site) args)
        Nothing -> callStmt (hsName site) args

-- ---------------------------------------------------------------------------
-- The offload predicate

-- | A site builder (given the site's name) when the statement is a loop the
-- device runs as the printed C does.
offloadableLoop :: TypeEnv -> [Stmt] -> Stmt -> Maybe (CVar -> HipSite)
offloadableLoop env whole s@(SLoop spec body) = do
  (kspec, kbody, macc) <- kernelLoop spec body
  guard' (all (supportedStmt False) kbody)
  guard' (scatterOk kspec kbody)
  let assigned = S.fromList (assignedVars [s])
      used = S.fromList (usedVars [s])
      free = S.toList ((used `S.difference` assigned) `S.union` maybe S.empty S.singleton macc)
      outside = S.fromList (usedVars (removeStmt s whole))
      liveOut = (assigned `S.intersection` outside) `S.difference` maybe S.empty S.singleton macc
  guard' (S.null liveOut)
  args <- mapM (\v -> (,) v <$> argType v) (sort free)
  accT <- case macc of
    Nothing -> Just Nothing
    Just acc -> (\t -> Just (acc, t)) <$> scalarType acc
  let vars = S.toList (assigned `S.union` used)
      types = M.fromList [(v, t) | v <- vars, Just t <- [M.lookup v env], t /= CTUnknown]
  pure $ \name -> HipSite {hsName = name, hsArgs = args, hsResult = accT, hsTypes = types, hsLoop = s}
  where
    argType v = case M.lookup v env of
      Just t@(CTArray e) | e `elem` [CTInt64, CTDouble, CTBool] -> Just t
      Just t | t `elem` [CTInt64, CTDouble, CTBool, CTTuple] -> Just t
      _ -> Nothing
    scalarType v = case M.lookup v env of
      Just t | t `elem` [CTInt64, CTDouble, CTBool] -> Just t
      _ -> Nothing
offloadableLoop _ _ _ = Nothing

guard' :: Bool -> Maybe ()
guard' b = if b then Just () else Nothing

-- | A site builder for an iterate loop (Lowering.hs:840-857) whose whole
-- step the device runs: the loop is serial with one counter the step never
-- reads; the step ends with cur = next, allocates next once at its top
-- level and assigns it nowhere else, and otherwise holds scalar statements
-- of scalars (no element loads), branches, and map loops offloadable on
-- their own.  The site takes cur among its arguments and returns the final
-- state.
offloadableIterate :: TypeEnv -> [Stmt] -> Stmt -> Maybe (CVar -> HipSite)
offloadableIterate env whole s@(SLoop spec body) = do
  guard' (lsRole spec == LoopIterate && lsExec spec == Serial && isNothing (lsRed spec))
  [t] <- Just (lsIters spec)
  guard' (not (null body))
  let step = init body
  SAssign cur (RAtom (AVar nxt)) <- Just (last body)
  ct@(CTArray e) <- M.lookup cur env
  guard' (e `elem` [CTInt64, CTDouble, CTBool])
  guard' (length [() | SAssign v (RArrayAlloc _) <- step, v == nxt] == 1)
  guard' (length [() | SAssign v _ <- flatten step, v == nxt] == 1)
  guard' (cur `notElem` assignedVars step)
  guard' (t `notElem` usedVars step)
  guard' (all (stepStmt nxt) step)
  let assigned = S.fromList (assignedVars [s])
      used = S.fromList (usedVars [s])
      free = S.toList ((used `S.difference` assigned) `S.union` S.singleton cur)
      outside = S.fromList (usedVars (removeStmt s whole))
      liveOut = (assigned `S.intersection` outside) `S.difference` S.singleton cur
  guard' (S.null liveOut)
  args <- mapM (\v -> (,) v <$> argType v) (sort free)
  let vars = S.toList (assigned `S.union` used)
      types = M.fromList [(v, ty) | v <- vars, Just ty <- [M.lookup v env], ty /= CTUnknown]
  pure $ \name -> HipSite {hsName = name, hsArgs = args, hsResult = Just (cur, ct), hsTypes = types, hsLoop = s}
  where
    argType v = case M.lookup v env of
      Just ty@(CTArray el) | el `elem` [CTInt64, CTDouble, CTBool] -> Just ty
      Just ty | ty `elem` [CTInt64, CTDouble, CTBool, CTTuple] -> Just ty
      _ -> Nothing
    scalarVar v = case M.lookup v env of
      Just ty -> ty `elem` [CTInt64, CTDouble, CTBool, CTTuple]
      Nothing -> False
    noLoad r = case r of
      RArrayLoad _ _ -> False
      _ -> True
    stepStmt nxt st = case st of
      SAssign v (RArrayAlloc _) -> v == nxt
      SAssign v r -> scalarVar v && supportedRHS r && noLoad r
      SIf _ a b -> all (stepStmt nxt) a && all (stepStmt nxt) b
      SLoop _ _ -> case offloadableLoop env whole st of
        Just mk -> isNothing (hsResult (mk "_"))
        Nothing -> False
      _ -> False
offloadableIterate _ _ _ = Nothing

-- | The loop whose iterations the kernel distributes, its body, and the
-- reduction accumulator if any.
kernelLoop :: LoopSpec -> [Stmt] -> Maybe (LoopSpec, [Stmt], Maybe CVar)
kernelLoop spec body
  | lsRole spec == LoopReductionWrapper
  , [_] <- lsIters spec
  , [IConst 1] <- lsBounds spec
  , [SLoop inner ib] <- body
  , Just rs <- lsRed inner
  , [_] <- lsIters inner =
      Just (inner, ib, Just (rsAccVar rs))
  | Just rs <- lsRed spec, [_] <- lsIters spec = Just (spec, body, Just (rsAccVar rs))
  | Nothing <- lsRed spec, independent = Just (spec, body, Nothing)
  | otherwise = Nothing
  where
    independent = case lsExec spec of
      Parallel _ -> True
      _ -> lsRole spec `elem` [LoopMap, LoopMapReduction]

-- | A combining scatter strategy needs the (+) / (+.) combine.
scatterOk :: LoopSpec -> [Stmt] -> Bool
scatterOk spec body = case lsExec spec of
  Parallel ps | psStrategy ps `elem` [ParallelGeneric, ParallelScatterDirect] -> True
  Parallel _ -> length (lsIters spec) == 1 && hasAddCombine body
  _ -> True
  where
    hasAddCombine ss = any triple (tails3 ss) || any inIf ss
    inIf t = case t of
      SIf _ a b -> hasAddCombine a || hasAddCombine b
      _ -> False
    tails3 (a : b : c : rest) = (a, b, c) : tails3 (b : c : rest)
    tails3 _ = []
    triple (SAssign old (RArrayLoad (AVar d) r), SAssign new (RBinOp op x y), SArrayWrite (AVar d') r' (AVar new')) =
      op `elem` [CAdd, CAddF] && d == d' && r == r' && new == new' && ((x == AVar old) /= (y == AVar old))
    triple _ = False

supportedStmt :: Bool -> Stmt -> Bool
supportedStmt inLoop s = case s of
  SAssign _ r -> supportedRHS r
  SArrayWrite a i v -> all simpleAtom [a, i, v]
  SIf c t e -> simpleAtom c && all (supportedStmt inLoop) t && all (supportedStmt inLoop) e
  SLoop _ b -> all (supportedStmt True) b
  SBreak -> inLoop
  _ -> False

supportedRHS :: RHS -> Bool
supportedRHS r = all simpleAtom (rhsAtoms r) && case r of
  RAtom _ -> True
  RBinOp _ _ _ -> True
  RUnOp op _ -> op `notElem` [CLog, CSin, CCos, CErf]
  RArrayLoad _ _ -> True
  RTuple as -> length as <= 8
  RProj _ _ -> True
  RArrayShape _ -> True
  RShapeSize _ -> True
  RShapeInit _ -> True
  RShapeLast _ -> True
  RFlatToNd _ _ -> True
  RNdToFlat _ _ -> True
  R2DToFlat _ _ -> True
  _ -> False

simpleAtom :: Atom -> Bool
simpleAtom a = case a of
  AString _ -> False
  AVecVar _ -> False
  _ -> True

assignedVars :: [Stmt] -> [CVar]
assignedVars ss = concat [f t | t <- flatten ss]
  where
    f t = case t of
      SAssign v _ -> [v]
      SLoop spec _ -> lsIters spec
      _ -> []

usedVars :: [Stmt] -> [CVar]
usedVars ss = concat [f t | t <- flatten ss]
  where
    atomVars as = [v | AVar v <- as]
    f t = case t of
      SAssign _ r -> atomVars (rhsAtoms r)
      SArrayWrite a i v -> atomVars [a, i, v]
      SIf c _ _ -> atomVars [c]
      SReturn a -> atomVars [a]
      SLoop spec _ ->
        concatMap ixVars (lsBounds spec ++ lsOrigins spec) ++ maybe [] (ixVars . rsInit) (lsRed spec)
      _ -> []
    ixVars e = case e of
      IVar v -> [v]
      IConst _ -> []
      IAdd a b -> ixVars a ++ ixVars b
      ISub a b -> ixVars a ++ ixVars b
      IMul a b -> ixVars a ++ ixVars b
      IDiv a b -> ixVars a ++ ixVars b
      ITuple es -> concatMap ixVars es
      IProj _ x -> ixVars x
      IFlatToNd a b -> ixVars a ++ ixVars b
      INdToFlat a b -> ixVars a ++ ixVars b
      ICall _ es -> concatMap ixVars es

-- | The procedure body without one statement (for the live-out check).
removeStmt :: Stmt -> [Stmt] -> [Stmt]
removeStmt target = mapMaybe go
  where
    go t
      | t == target = Nothing
      | otherwise = Just $ case t of
          SLoop spec b -> SLoop spec (removeStmt target b)
          SIf c a b -> SIf c (removeStmt target a) (removeStmt target b)
          SParallelRegion b -> SParallelRegion (removeStmt target b)
          _ -> t
