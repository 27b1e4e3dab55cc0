"""Build the Fortran drivers (amdflang) for the test cases.

For each case the include files are emitted in the reference's generated
format (prep/fortran_emit.py) and gatt_driver.F90 is compiled against them and
linked with libgatt_dns.so.  Executables go to gatt_dns_amd/_lib/.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(HERE), "_lib")
FC = os.environ.get("FC_AMD", shutil.which("amdflang") or "/opt/rocm/bin/amdflang")

# case name -> builder (module-level so tests can rebuild the same inputs)
CASES = {
    "cavity33": lambda cases: cases.lid_driven(33),
    "box24": lambda cases: cases.turbulence_box(24),
}


def build(verbose=True):
    from gatt_dns_amd.prep import cases
    from gatt_dns_amd.prep.fortran_emit import write_includes
    if not os.path.exists(FC):
        raise RuntimeError("amdflang not found")
    os.makedirs(LIB_DIR, exist_ok=True)
    work = os.path.join(LIB_DIR, "fortran_build")
    os.makedirs(work, exist_ok=True)
    mod_obj = os.path.join(work, "gatt_hip.o")
    subprocess.run([FC, "-O2", "-c", os.path.join(HERE, "gatt_hip.f90"), "-module-dir", work, "-o", mod_obj],
                   check=True, capture_output=True, text=True)
    out = []
    for name, mk in CASES.items():
        inc = os.path.join(work, name)
        write_includes(mk(cases), inc)
        exe = os.path.join(LIB_DIR, "gatt_driver_%s" % name)
        cmd = [FC, "-O2", "-cpp", "-I", inc, "-I", work, os.path.join(HERE, "gatt_driver.F90"), mod_obj,
               "-L", LIB_DIR, "-lgatt_flowio", "-lgatt_dns", "-Wl,-rpath,$ORIGIN", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed for %s:\n%s" % (name, r.stderr[-4000:]))
        out.append(exe)
        if verbose:
            print("built", exe)
    return out


if __name__ == "__main__":
    build()
