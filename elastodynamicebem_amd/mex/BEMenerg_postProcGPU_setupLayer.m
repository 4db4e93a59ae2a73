function campoVett = BEMenerg_postProcGPU_setupLayer(pbParam, domainMesh, density, numPoints, xVal, tVal, methodInfo, PPn, PPw)
% Drop-in for functions/BEMenerg_indir/BEMenerg_postProc/BEMenerg_postProcGPU/BEMenerg_postProcGPU_setupLayer.m:
% same signature and output (xSize x tSize cell of 3 x 1), the numeric GPU
% post-processing (kernelLayer.cu) on the MI355X through ebem_mex.  PPn/PPw are
% implied by methodInfo.numSubRegionPP / numNodiSingPP.
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
ebem_mex('pp_rule', methodInfo.numSubRegionPP, methodInfo.numNodiSingPP);
u = ebem_mex('pp_numeric', density, xVal, tVal);
xSize = size(xVal, 1);
tSize = length(tVal);
if numPoints ~= xSize * tSize
    error('numPoints does not match xVal x tVal')
end
campoVett = cell(xSize, tSize);
for k = 1 : tSize
    for p = 1 : xSize
        campoVett{p, k} = u(:, p, k);
    end
end
end
