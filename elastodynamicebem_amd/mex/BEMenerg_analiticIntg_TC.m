function [ris1, ris2, ris3, ris4] = BEMenerg_analiticIntg_TC(zeta, RS, rhoS, THETA1, THETA2, infoTri2D, B)
% Drop-in for functions/BEMenerg_analiticIntg/BEMenerg_analiticIntg_TC.m
% (P region between the S circle and the edge): the same closed form, evaluated on the MI355X by the
% library's region routine (ebem_mex('init') must have run).  infoTri2D is the
% struct of calcoloDatiTriangolo.m, B the 3x3x6 array of coeffB.m.
info = [infoTri2D.a, infoTri2D.b, infoTri2D.c, infoTri2D.cos_alpha, infoTri2D.cos_beta, ...
        infoTri2D.cos_gamma, infoTri2D.sin_alpha, infoTri2D.sin_beta, infoTri2D.sin_gamma, ...
        infoTri2D.alpha, infoTri2D.beta, infoTri2D.gamma];
[ris1, ris2, ris3, ris4] = ebem_mex('intg', 3, [RS, rhoS, 0, 0, zeta, THETA1, THETA2], info, B);
end
