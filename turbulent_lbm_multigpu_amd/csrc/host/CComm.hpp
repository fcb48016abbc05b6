// CComm.hpp -- one face exchange seen from the receiving subdomain
// (src/CComm.hpp:8-79): neighbour id, send/recv boxes and the inward normal.
#pragma once

#include "CVector.hpp"

template <typename T>
class CComm {
    int _dstID;
    CVector<3, int> _send_size, _recv_size, _send_origin, _recv_origin, _comm_direction;

public:
    CComm(int dstID, CVector<3, int> send_size, CVector<3, int> recv_size, CVector<3, int> send_origin,
          CVector<3, int> recv_origin, CVector<3, int> comm_direction)
        : _dstID(dstID), _send_size(send_size), _recv_size(recv_size), _send_origin(send_origin),
          _recv_origin(recv_origin), _comm_direction(comm_direction) {}

    CVector<3, int> getCommDirection() const { return _comm_direction; }
    void setCommDirection(CVector<3, int> n) { _comm_direction = n; }
    CVector<3, int> getRecvOrigin() const { return _recv_origin; }
    void setRecvOrigin(CVector<3, int> o) { _recv_origin = o; }
    CVector<3, int> getRecvSize() const { return _recv_size; }
    void setRecvSize(CVector<3, int> s) { _recv_size = s; }
    CVector<3, int> getSendOrigin() const { return _send_origin; }
    void setSendOrigin(CVector<3, int> o) { _send_origin = o; }
    CVector<3, int> getSendSize() const { return _send_size; }
    void setSendSize(CVector<3, int> s) { _send_size = s; }
    int getDstId() const { return _dstID; }
    void setDstId(int id) { _dstID = id; }
};
