function campoVett = BEMenerg_ppD_setup(pbParam, domainMesh, density, numPoints, xVal, tVal, methodInfo, PPn, PPw)
% Drop-in for functions/BEMenerg_dirPP/BEMenerg_ppD_setup.m: the direct
% post-processing u = sum_y G_y alpha_y - P_y g_y (BEMenerg_ppD_calc.m) at every
% point of xVal and time of tVal, with the domain's Dirichlet datum
% (BEMenerg_dir_calcBoundDataDirichlet.m).  Returns the xSize x tSize cell of
% 3x1 vectors and saves it as the reference does.  PPn / PPw are implied by
% methodInfo.numSubRegionPP x methodInfo.numNodiSingPP (the rule the reference
% builds them from).
xSize = size(xVal, 1);
tSize = length(tVal);
if numPoints ~= xSize * tSize
    error('numPoints does not match xVal x tVal')
end
switch pbParam.domainType
    case 'barH1'
        datum = 4;
    case 'barH3'
        datum = 5;
    case 'sphereWave'
        datum = 6;
    case 'elementoIndustriale'
        datum = 7;
    otherwise
        error('Problema non codificato')   % calcBoundDataDirichlet.m has no datum for it
end
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
ebem_mex('pp_rule', methodInfo.numSubRegionPP, methodInfo.numNodiSingPP);
gK = ebem_mex('dirichlet_datum', datum);
u = ebem_mex('pp_direct', density, gK, xVal, tVal);
campoVett = cell(xSize, tSize);
for ix = 1:xSize
    for it = 1:tSize
        campoVett{ix, it} = u(:, ix, it);
    end
end
filePath = strcat("./outputData/", pbParam.domainType, num2str(pbParam.lev), "_NEW_campoVett");
save(filePath, 'campoVett');
end
