function density = BEMenerg_coreSA_timeMarching(pbParam, domainMesh, methodInfo, GHn, GHw)
% Drop-in for functions/BEMenerg_indir/BEMenerg_core/BEMenerg_coreSA/BEMenerg_coreSA_timeMarching.m:
% same signature and output (3N x nT density); assembly, RHS and MOT run on
% the MI355X through ebem_mex.  GHn/GHw are implied by methodInfo.numNodiExt.
N = domainMesh.numberTriangles;
ebem_mex('init', 0);
ebem_mex('mesh', domainMesh.coordinates, domainMesh.triangles(:, 1:3));
ebem_mex('problem', pbParam.rho, pbParam.mu, pbParam.lambda, pbParam.Tfin, pbParam.nT);
ebem_mex('rule', methodInfo.numNodiExt);
ebem_mex('assemble', 0);
switch pbParam.domainType
    case {'sphereUniform', 'sphereNotUniform'}
        datum = 2;
    case {'screenUniform', 'screenGraded'}
        datum = 3;
    otherwise
        error('Caso ancora non riportato/Errore nei dati')
end
ebem_mex('rhs', datum);
density = ebem_mex('mot', []);
end
