function density = BEMenerg_dir_timeMarchingN(pbParam, domainMesh, methodInfo, GHn, GHw, GHCn, GHCw, DIAGn, DIAGw)
% Drop-in for functions/BEMenerg_dir/BEMenerg_dir_timeMarchingN.m: same
% signature and output (3Nv x nT density at the vertices) for "FN nExt nSub
% nSing nDiag" methods; V, K, betaV, K_0 + I_Gamma and the least-squares
% marching run on the MI355X through ebem_mex.
if (pbParam.lambda + pbParam.mu == 0)
    error("Problema non adatto a questa formulazione")
end
switch pbParam.domainType
    case 'barH1'
        datum = 4;
    case 'barH3'
        datum = 5;
    otherwise
        error('Problema non codificato')
end
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
ebem_mex('rule', methodInfo.numNodiExt);
ebem_mex('inner_rule', methodInfo.numSubRegion, methodInfo.numNodiSing, methodInfo.numNodiDiag);
ebem_mex('assemble', 4);
ebem_mex('assemble_k');
density = ebem_mex('mot_neumann', datum);
end
