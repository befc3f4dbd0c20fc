-- |
-- Module: HIPBackend
--
-- Compiles and runs what Language.Hydrangea.CodegenHIP printed: the same
-- flow as the reference CLI's C backend (app/CBackend.hs:78-106, compiler
-- chosen by --cc, then $CC, then cc), with the program linked against
-- libhydhip instead of compiled with the runtime's C file and OpenMP.
-- HYD_HIP_HOME names the root of the MI355X backend (include/ and
-- hydrangea-compiler_amd/lib/libhydhip.so).
--
-- UNTESTED: GHC is not available where this module was written.  The
-- compile line it builds is the one the repository's tests use for their
-- gcc-built hosts.
module HIPBackend (compileAndRunHIP, hipCompileArgs) where

import Control.Monad (when)
import System.Directory (copyFile)
import System.Environment (lookupEnv)
import System.Exit (ExitCode (..))
import System.FilePath ((</>))
import System.IO (hPutStrLn, stderr)
import System.IO.Temp (withSystemTempDirectory)
import System.Process (createProcess, proc, readProcessWithExitCode, waitForProcess)

-- | Compiler arguments for one generated C file.
hipCompileArgs :: FilePath -> FilePath -> FilePath -> [String]
hipCompileArgs home cFile exe =
  [ "-std=c99", "-O3", "-I" ++ (home </> "include"), cFile, "-o", exe
  , "-L" ++ libDir, "-lhydhip", "-Wl,-rpath," ++ libDir, "-lm"
  ]
  where
    libDir = home </> "hydrangea-compiler_amd" </> "lib"

-- | Writes the source, compiles it, runs it; keepC copies the source to
-- the working directory as the C backend's --keep-c does.
compileAndRunHIP :: Maybe String -> Bool -> String -> IO ExitCode
compileAndRunHIP mcc keepC src = do
  mhome <- lookupEnv "HYD_HIP_HOME"
  envCC <- lookupEnv "CC"
  case mhome of
    Nothing -> do
      hPutStrLn stderr "HIPBackend: set HYD_HIP_HOME to the MI355X backend's root"
      pure (ExitFailure 1)
    Just home -> withSystemTempDirectory "hydrangea-hip" $ \dir -> do
      let cFile = dir </> "program.c"
          exe = dir </> "program"
          cc = pick mcc (pick envCC "cc")
      writeFile cFile src
      when keepC (copyFile cFile "program.c")
      (code, _, err) <- readProcessWithExitCode cc (hipCompileArgs home cFile exe) ""
      case code of
        ExitFailure _ -> hPutStrLn stderr err >> pure code
        ExitSuccess -> do
          (_, _, _, ph) <- createProcess (proc exe [])
          waitForProcess ph
  where
    pick (Just s) _ | not (null s) = s
    pick _ d = d
