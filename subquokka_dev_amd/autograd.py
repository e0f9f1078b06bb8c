"""torch.autograd.Function over the engine's adjoint gradient.

    energy = HEAEnergy.apply(theta, engine, layers, checkpoint_every)
    energy = CircuitEnergy.apply(theta, engine, text, param, terms, checkpoint_every)

CircuitEnergy: any parameterised circuit (circuit text, param[i] = index of
gate i's angle in theta or -1) from |0...0>, observable a sum of Z strings
[(coef, qubits), ...]; one sq_adjoint_grad call gives the energy and the whole
gradient.

theta is a 1-D float64 tensor of 2 * n * layers angles (RY angles then RZ
angles, layer by layer); engine is a subquokka_dev_amd.Engine of n qubits.
The forward pass runs the ansatz and the reverse sweep on the GPU through
sq_adjoint_grad_hea (one call computes the energy and the whole gradient);
backward returns grad_output times the stored gradient.  PyTorch only carries
the 2nL angles: the state never leaves the engine.
"""
import numpy as np
import torch


class HEAEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, layers, checkpoint_every=0):
        th = theta.detach().to(torch.float64).cpu().numpy()
        energy, grad = engine.adjoint_grad_hea(layers, th, checkpoint_every)
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(grad)).to(theta.device, theta.dtype))
        return torch.tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output * grad, None, None, None


def hea_energy(theta, engine, layers, checkpoint_every=0):
    return HEAEnergy.apply(theta, engine, layers, checkpoint_every)


class CircuitEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, text, param, terms, checkpoint_every=0):
        th = theta.detach().to(torch.float64).cpu().numpy()
        engine.init_zero()
        energy, grad = engine.adjoint_grad(text, param, th, terms, checkpoint_every)
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(grad)).to(theta.device, theta.dtype))
        return torch.tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output * grad, None, None, None, None, None


def circuit_energy(theta, engine, text, param, terms, checkpoint_every=0):
    return CircuitEnergy.apply(theta, engine, text, param, terms, checkpoint_every)
