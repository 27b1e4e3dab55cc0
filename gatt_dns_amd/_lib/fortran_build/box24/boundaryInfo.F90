    nUd = 0
    allocate(iUd(1,6))
    allocate(vUd(1))
    nVd = 0
    allocate(iVd(1,6))
    allocate(vVd(1))
    nWd = 0
    allocate(iWd(1,6))
    allocate(vWd(1))
    nEd = 0
    allocate(iEd(1,6))
    allocate(vEd(1))
    nPd = 0
    allocate(iPd(1,6))
    allocate(vPd(1))
    nUn = 0
    allocate(iUn(1,6))
    allocate(dUn(1))
    nVn = 0
    allocate(iVn(1,6))
    allocate(dVn(1))
    nWn = 0
    allocate(iWn(1,6))
    allocate(dWn(1))
    nEn = 0
    allocate(iEn(1,6))
    allocate(dEn(1))
    nPn = 0
    allocate(iPn(1,6))
    allocate(dPn(1))
    nUs = 0
    allocate(iUs(1,6))
    allocate(dUs(1))
    nVs = 0
    allocate(iVs(1,6))
    allocate(dVs(1))
    nWs = 0
    allocate(iWs(1,6))
    allocate(dWs(1))
    nEs = 0
    allocate(iEs(1,6))
    allocate(dEs(1))
    nPs = 0
    allocate(iPs(1,6))
    allocate(dPs(1))
    cN = 0
    allocate(cL(1,6))
    allocate(cD(1,3))
    allocate(cAdiabatic(1))
