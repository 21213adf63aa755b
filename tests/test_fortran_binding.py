"""The iso_c_binding module the Fortran host uses (mulskips_amd/fortran/
mskg_bindings.f90) against the C header: both sides must agree on the layout
of mskg_config and mskg_event, and a Fortran program must link against
libmskg.so and reach mskg_create (which reports MSKG_EDEVICE without a GPU,
or MSKG_OK with one).  Needs ROCm's flang and gcc; skipped without them."""
import os
import shutil
import subprocess

import pytest

from tests.helpers import ROOT

FLANG = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
LIB = os.path.join(ROOT, "mulskips_amd", "libmskg.so")

C_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "mskg.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(mskg_config), sizeof(mskg_event), offsetof(mskg_config, seed),
         offsetof(mskg_config, second_factor), offsetof(mskg_config, reserved));
  return 0;
}
"""

F_PROBE = r"""
program probe
  use iso_c_binding
  use mskg_bindings
  implicit none
  type(mskg_config), target :: cfg
  type(mskg_event), target :: ev(2)
  type(c_ptr) :: eng
  real(c_double), target :: pe(12288), pd(9)
  integer(c_int) :: rc
  write(*, '(5I8)') c_sizeof(cfg), c_sizeof(ev(1)), &
      loc_diff(c_loc(cfg%seed)), loc_diff(c_loc(cfg%second_factor)), loc_diff(c_loc(cfg%reserved))
  pe = 0; pd = 1
  cfg%len_x = 96; cfg%len_y = 96; cfg%len_z = 144
  cfg%n_crystal = 1; cfg%n_cov = 0
  cfg%list_crystal = [14, 0, 0]; cfg%list_cov = 0
  cfg%ptrans_e = c_loc(pe); cfg%ptrans_d = c_loc(pd)
  cfg%ptrans_zig = 1.0d0; cfg%la_mode = 0; cfg%seed = 1
  cfg%tile = 0; cfg%tile_capacity = 0; cfg%events_per_pass = 0; cfg%tau_policy = 0; cfg%frozen_below_z = 0
  cfg%reserved = 0
  rc = mskg_create(cfg, eng)
  write(*, '(I4)') rc
  if (rc /= MSKG_OK) write(*, '(A)') trim(mskg_error_text(eng))
  call mskg_destroy(eng)
contains
  integer function loc_diff(p)
    type(c_ptr), intent(in) :: p
    integer(c_intptr_t) :: a, b
    a = transfer(p, a)
    b = transfer(c_loc(cfg), b)
    loc_diff = int(a - b)
  end function
end program probe
"""


@pytest.mark.skipif(not FLANG or not shutil.which("gcc"), reason="needs amdflang and gcc")
def test_fortran_binding_layout_and_link(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libmskg.so not built")
    (tmp_path / "probe.c").write_text(C_PROBE)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(tmp_path / "cprobe"),
                    str(tmp_path / "probe.c")], check=True)
    c_out = subprocess.run([str(tmp_path / "cprobe")], capture_output=True, text=True, check=True).stdout.split()
    (tmp_path / "probe.f90").write_text(F_PROBE)
    subprocess.run([FLANG, "-J", str(tmp_path), "-o", str(tmp_path / "fprobe"),
                    os.path.join(ROOT, "mulskips_amd", "fortran", "mskg_bindings.f90"), str(tmp_path / "probe.f90"),
                    LIB, f"-Wl,-rpath,{os.path.dirname(LIB)}"], check=True, cwd=tmp_path)
    res = subprocess.run([str(tmp_path / "fprobe")], capture_output=True, text=True, timeout=120)
    lines = res.stdout.split("\n")
    f_layout = lines[0].split()
    assert f_layout == c_out, (f_layout, c_out)
    rc = int(lines[1])
    assert rc in (0, 4), res.stdout  # MSKG_OK on a GPU, MSKG_EDEVICE without one
    if rc == 4:
        assert "no HIP device" in res.stdout
