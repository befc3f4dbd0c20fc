-- |
-- Module: HIPBackend
--
-- The CLI's MI355X backend, beside app/CBackend.hs: it writes what
-- Language.Hydrangea.CodegenHIP produced - the printed C and the offloaded
-- loops, one HipSite per line - to a temporary directory and hands both to
-- hydhip-cc ($HYD_HIP_HOME/hydrangea-compiler_amd/hydhip-cc), which runs the
-- kernel generator (hydgen), compiles the kernels with
-- `hipcc --offload-arch=gfx950 -O3 -ffp-contract=off`, compiles the C with
-- the chosen compiler as the C backend does (-O3 -march=native -std=c99, and
-- -fopenmp after the same probe, app/CBackend.hs:29-49, for the loops that
-- stay on the host), and links libhydhip.  Then, as compileAndRunC does
-- (app/CBackend.hs:78-106), it keeps the artifacts if asked and runs the
-- program unless compile-only.  The compiler is chosen in the C backend's
-- order: --cc, then $CC, then cc.
--
-- UNTESTED here: GHC is not available where this module was written.  The
-- command it runs is tested: tests/test_hydhip_cc.py builds programs with
-- exactly these arguments, and tests/test_gpu_hydhip_cc.py runs them on the
-- GPU against the oracle.
module HIPBackend (compileAndRunHIP, hydhipCcArgs) where

import Control.Monad (when)
import Language.Hydrangea.CodegenHIP (HIPArtifacts (..), renderSites)
import System.Directory (copyFile)
import System.Environment (lookupEnv)
import System.Exit (ExitCode (..))
import System.FilePath ((</>))
import System.IO (hPutStrLn, stderr)
import System.IO.Temp (withSystemTempDirectory)
import System.Process (createProcess, proc, readProcessWithExitCode, waitForProcess)

-- | The arguments of hydhip-cc for one program.
hydhipCcArgs :: Maybe String -> Bool -> FilePath -> FilePath -> FilePath -> [String]
hydhipCcArgs mcc parallel cFile sitesFile exe =
  [cFile, sitesFile, "-o", exe] ++ maybe [] (\cc -> ["--cc", cc]) mcc ++ ["--no-openmp" | not parallel]

-- | Compile the program and (unless compileOnly) run it.  keepC copies the
-- C, the sites and the executable to the working directory as
-- hydrangea_out.c, hydrangea_out.sites and hydrangea_out.
compileAndRunHIP :: Maybe String -> HIPArtifacts -> Bool -> Bool -> Bool -> IO ExitCode
compileAndRunHIP mcc arts keepC compileOnly parallel = do
  mhome <- lookupEnv "HYD_HIP_HOME"
  envCC <- lookupEnv "CC"
  case mhome of
    Nothing -> do
      hPutStrLn stderr "HIPBackend: set HYD_HIP_HOME to the MI355X backend's root"
      pure (ExitFailure 1)
    Just home -> withSystemTempDirectory "hydrangea-hip" $ \dir -> do
      let cFile = dir </> "out.c"
          sitesFile = dir </> "out.sites"
          exe = dir </> "hydrangea_out"
          driver = home </> "hydrangea-compiler_amd" </> "hydhip-cc"
          cc = pick mcc envCC
      writeFile cFile (hipC arts)
      writeFile sitesFile (renderSites (hipSites arts) (hipGroups arts))
      (code, _, err) <- readProcessWithExitCode driver (hydhipCcArgs cc parallel cFile sitesFile exe) ""
      case code of
        ExitFailure _ -> putStrLn "HIP compilation failed:" >> putStrLn err >> pure code
        ExitSuccess -> do
          when keepC $ do
            copyFile cFile "hydrangea_out.c"
            copyFile sitesFile "hydrangea_out.sites"
            copyFile exe "hydrangea_out"
          if compileOnly
            then pure ExitSuccess
            else do
              (_, _, _, ph) <- createProcess (proc exe [])
              waitForProcess ph
  where
    pick (Just s) _ | not (null s) = Just s
    pick _ (Just s) | not (null s) = Just s
    pick _ _ = Nothing
