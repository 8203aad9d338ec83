! dropin_stubs.f90 -- TEST STAND-INS for the host LES's velocity boundary
! conditions (src_v70/bounds.for: boundu, boundv, boundw), which newsolv_mg
! calls after its loop (newsolv_mg.for:78-80).  They do nothing here: the
! drop-in and the reference's loop replay are compared on the state the
! loop itself leaves.  Never linked into the product.
subroutine boundu
end subroutine boundu

subroutine boundv
end subroutine boundv

subroutine boundw
end subroutine boundw
