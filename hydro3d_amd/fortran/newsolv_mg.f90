! newsolv_mg.f90 -- the name the LES driver calls (src_v70/flosol.for:215:
! `call newsolv_mg`), bound to the device path.  Build the LES with this
! object (and newsolv_mg_h3d.o, h3dmg_mod.o, -lh3dmg) in place of
! newsolv_mg.o (src_v70/makefile's object list): nothing else changes.
subroutine newsolv_mg
  call newsolv_mg_h3d
end subroutine newsolv_mg
