    nDist = 0
    allocate(dist(max(nDist,1)))
    allocate(distX(max(0,1)))
    nForc = 1
    allocate(forc(max(nForc,1)))
    forc(1)%kind = 1
    forc(1)%box = (/1,24,1,24,1,24/)
    forc(1)%par(1) = 1.00000000000000000000d0
    forc(1)%par(2) = 19.84126984126984893919d0
