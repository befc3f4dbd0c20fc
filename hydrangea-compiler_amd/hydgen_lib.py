"""ctypes bindings of a library of hydgen-generated launchers.

hydgen writes, beside the HIP source of a program's offloaded loops, one line
per launcher: `name result : type:arg type:arg ...` with each type one of
i64, f64, tuple (hyd_tuple_t by value) and array (hyd_array_t*).  Program()
loads the compiled library and declares every launcher from those lines, so
a Python host calls them exactly as the printed C does, positionally or by
argument name (Program.call).  Arrays are the library's
hyd_array_t (hydhip.host_array / hyd_array_alloc_device); a launcher leaves
the arrays it writes in HBM, and hyd_array_sync_host brings them back.
"""
import ctypes
import os

import hydhip

HERE = os.path.dirname(os.path.abspath(__file__))
PROGRAMS_LIB = os.path.join(HERE, "lib", "libhydgen_programs.so")
PROGRAMS_SIG = os.path.join(HERE, "lib", "libhydgen_programs.sig")

_TYPES = {
    "i64": ctypes.c_int64,
    "f64": ctypes.c_double,
    "tuple": hydhip.HydTuple,
    "array": ctypes.POINTER(hydhip.HydArray),
}


def parse_signatures(text):
    """{name: (result type, [(argument type, argument name)])} from the
    signature lines."""
    out = {}
    for line in text.splitlines():
        if not line.strip():
            continue
        head, _, args = line.partition(" : ")
        name, res = head.split()
        out[name] = (res, [tuple(a.split(":", 1)) for a in args.split()])
    return out


class Program:
    def __init__(self, so_path=PROGRAMS_LIB, sig_path=PROGRAMS_SIG):
        if not os.path.exists(so_path) or not os.path.exists(sig_path):
            raise hydhip.HydError(hydhip.HYD_E_NODEV,
                                  "%s not built: run hydrangea-compiler_amd/build.py" % so_path)
        hydhip.lib()  # libhydhip first: the launchers call into it
        self._cdll = ctypes.CDLL(so_path)
        with open(sig_path) as f:
            self.signatures = parse_signatures(f.read())
        self._fns = {}
        for name, (res, args) in self.signatures.items():
            fn = getattr(self._cdll, name)
            fn.restype = _TYPES[res]
            fn.argtypes = [_TYPES[t] for t, _ in args]
            self._fns[name] = fn

    def call(self, name, env):
        """Launcher `name` with its arguments taken from env by name (tuples
        as Python sequences of ints)."""
        args = []
        for t, a in self.signatures[name][1]:
            x = env[a]
            args.append(hydhip.hyd_tuple(*x) if t == "tuple" else x)
        return self._fns[name](*args)

    def __getattr__(self, name):
        fns = self.__dict__.get("_fns", {})
        if name not in fns:
            raise AttributeError("no generated launcher %s" % name)
        return fns[name]

    def launchers(self):
        return sorted(self._fns)


_default = None


def programs():
    """The example programs' launchers (lib/libhydgen_programs.so)."""
    global _default
    if _default is None:
        _default = Program()
    return _default
