function density = BEMenerg_dir_timeMarchingD(pbParam, domainMesh, methodInfo, GHn, GHw, GHCn, GHCw, DIAGn, DIAGw)
% Drop-in for functions/BEMenerg_dir/BEMenerg_dir_timeMarchingD.m: same
% signature and output (3N x nT density) for "FN nExt nSub nSing nDiag"
% methods; V (kernelV.cu form), K (kernelK.cu + calcSubBlockSingK.m),
% rhs/2 + betaK and the MOT run on the MI355X through ebem_mex.  The rules
% are implied by methodInfo.
if (pbParam.lambda + pbParam.mu == 0)
    error("Problema non adatto a questa formulazione")
end
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
ebem_mex('rule', methodInfo.numNodiExt);
ebem_mex('inner_rule', methodInfo.numSubRegion, methodInfo.numNodiSing, methodInfo.numNodiDiag);
ebem_mex('assemble', 4);
ebem_mex('assemble_k');
switch pbParam.domainType
    case {'sphereUniform', 'sphereNotUniform'}
        datum = 2;
    case {'screenUniform', 'screenGraded'}
        datum = 3;
    case 'barH1'
        datum = 4;
    case 'barH3'
        datum = 5;
    case 'sphereWave'
        datum = 6;
    case 'elementoIndustriale'
        datum = 7;
    otherwise
        error('Problema non codificato')
end
if datum < 4
    error('Problema non codificato')   % calcBoundDataDirichlet.m has no trace for it
end
ebem_mex('rhs_dirichlet', datum);
density = ebem_mex('mot', []);
end
