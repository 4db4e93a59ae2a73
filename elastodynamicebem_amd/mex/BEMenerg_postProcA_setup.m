function campoVett = BEMenerg_postProcA_setup(pbParam, domainMesh, density, numPoints, xVal, tVal)
% Drop-in for functions/BEMenerg_indir/BEMenerg_postProc/BEMenerg_postProcA/BEMenerg_postProcA_setup.m:
% returns the xSize x tSize cell of 3x1 displacement vectors.
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
u = ebem_mex('pp', density, xVal, tVal);
xSize = size(xVal, 1); tSize = length(tVal);
if numPoints ~= xSize * tSize
    error('numPoints does not match xVal x tVal')
end
campoVett = cell(xSize, tSize);
for ix = 1:xSize
    for it = 1:tSize
        campoVett{ix, it} = u(:, ix, it);
    end
end
end
